/*
 * tos_amd.h -- C-ABI of the MI355X (gfx950) pair-counting correlation
 * estimator: the drop-in boundary under turbo-octo-spice's C++ API.
 *
 * The reference exposes C++ classes only (no plugin system).  These entry
 * points are what its XiEstimator / SkyBins / HDF5Delta / bruteGPU call sites
 * would bind through a thin FFI; each declaration cites the reference
 * interface it replaces.  The C++ mirror of the reference classes
 * (turbo-octo-spice_amd/turbooctospice/) is implemented on top of them.
 *
 * Conventions
 *   - Every function returns an int status: TOS_OK (0) or a TOS_ERR_* code.
 *     Nothing throws across this boundary and nothing calls exit().
 *     tos_last_error(ctx) returns a message for the last failure on ctx.
 *   - Host arrays passed in are copied; the library never keeps a host
 *     pointer after a call returns.  Outputs go to caller-owned arrays.
 *   - One tos_ctx owns one HIP device, its streams and device buffers.
 *     A ctx is driven by one host thread at a time, except the job-queue
 *     calls (tos_job_*), which are thread safe.
 *   - All angles are radians.  Distances are comoving Mpc/h.
 */
#ifndef TOS_AMD_H
#define TOS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOS_ABI_VERSION 1

enum tos_status {
    TOS_OK = 0,
    TOS_ERR_INVALID = 1,     /* bad argument or inconsistent input */
    TOS_ERR_NOMEM = 2,       /* host or device allocation failed / over budget */
    TOS_ERR_DEVICE = 3,      /* HIP runtime error or no usable device */
    TOS_ERR_INTERNAL = 4,
    TOS_ERR_NOT_READY = 5,   /* tos_job_poll: job still running */
    TOS_ERR_CANCELLED = 6,   /* job was cancelled or its context shut down */
    TOS_ERR_TIMEOUT = 7      /* tos_job_wait deadline passed */
};

typedef struct tos_ctx tos_ctx;
typedef struct tos_catalog tos_catalog;
typedef struct tos_pairs tos_pairs;

/* ---------------------------------------------------------------------- */
/* Context                                                                 */
/* ---------------------------------------------------------------------- */

/* Replaces the implicit device set-up of gpu/brute.cu:98-131 and the
 * ThreadPool construction of XiEstimator.cc:32. */
int tos_ctx_create(int device, tos_ctx **out);
int tos_ctx_destroy(tos_ctx *ctx);
/* Replaces HANDLE_ERROR's printf+exit (gpu/brute.h:12-18) and the
 * tos::RuntimeError messages (RuntimeError.h:8-18). */
const char *tos_last_error(const tos_ctx *ctx);
/* Device description (gpu/brute.cu:121-149, src/deviceQuery.cu:6-37). */
typedef struct {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t wavefront_size;
    int64_t lds_per_block;
    int64_t global_mem;
    int32_t clock_khz;
} tos_device_info;
int tos_device_info_get(tos_ctx *ctx, tos_device_info *out);
const char *tos_build_info(void);

/* ---------------------------------------------------------------------- */
/* Binning grid                                                            */
/* ---------------------------------------------------------------------- */

/* XiEstimator::BinningCoordinateType (XiEstimator.h:19-21), same order. */
enum tos_coord { TOS_POLAR = 0, TOS_CARTESIAN = 1, TOS_OBSERVING = 2 };

/* AbsTwoPointGrid's per-axis parameters (AbsTwoPointGrid.h:55-58,71-85):
 * axis 0 = r_par | r | dloglam, axis 1 = r_perp | mu | theta,
 * axis 2 = log10(lya(1+z)).  max[k] is the high edge of the last bin. */
typedef struct {
    int32_t type;          /* enum tos_coord */
    int32_t nbins[3];
    double min[3];
    double width[3];
    double max[3];
} tos_grid;

/* likely::createBinning("[min:max]*n") as used by h5healxi.cc:138-140:
 * fills axis `axis` of *g.  Only uniform binnings are accepted. */
int tos_grid_parse_axis(const char *spec, tos_grid *g, int axis);
/* AbsTwoPointGrid::maxAngularScale (CartesianGrid.h:46, PolarGrid.h:46,
 * QuasarGrid.h:45). */
double tos_grid_max_angular_scale(const tos_grid *g, double scale);
/* AbsTwoPointGrid::getBinCenters (AbsTwoPointGrid.h:33,71-73). */
int tos_grid_bin_centers(const tos_grid *g, int32_t index, double centers[3]);

/* ---------------------------------------------------------------------- */
/* Sightline catalogue (device SoA)                                        */
/* ---------------------------------------------------------------------- */

/* The pixels of std::vector<Forest> (types.h:43-66) in SoA form: LOS k owns
 * pixels [offsets[k], offsets[k+1]).  Replaces the by-value Forest copies
 * handed to XiEstimator (XiEstimator.h:28-29,75). */
typedef struct {
    int64_t n_los;
    const int64_t *offsets;          /* n_los + 1 */
    const double *ra, *dec;          /* radians */
    const float *value, *loglam, *weight, *distance;
} tos_sightlines;

int tos_catalog_upload(tos_ctx *ctx, const tos_sightlines *s, tos_catalog **out);
int tos_catalog_free(tos_ctx *ctx, tos_catalog *cat);
int tos_catalog_info(const tos_catalog *cat, int64_t *n_los, int64_t *n_pix,
                     int64_t *device_bytes);

/* ---------------------------------------------------------------------- */
/* Sky binning and candidate sightline pairs (device CSR)                  */
/* ---------------------------------------------------------------------- */

/* SkyBins implementations: HealpixBins (HealpixBins.h:18-82, RING scheme,
 * nside = 2^order) and PlateBins (PlateBins.h:18-92). */
enum tos_sky_kind { TOS_SKY_HEALPIX = 0, TOS_SKY_PLATES = 1 };
typedef struct {
    int32_t kind;
    int32_t order;                   /* HEALPix order */
    /* PlateBins: plate of each LOS and the plate list (ra/dec radians) */
    const int32_t *plate_of_los;
    int64_t n_plates;
    const int32_t *plate_id;
    const double *plate_ra, *plate_dec;
} tos_skybins;

/* HealpixBins::ang2pix with radec2pnt (HealpixBins.h:58-72). */
int tos_healpix_radec2pix(int32_t order, int64_t n, const double *ra,
                          const double *dec, int32_t *out_pix);

/* Candidate enumeration of XiEstimator::skybin_xi_task
 * (XiEstimator.cc:121-159): for each occupied sky bin, each primary i in
 * it, each neighbouring bin within max_ang, each j > i in that bin, keep
 * the pair iff cos(sep) > cos(max_ang).  The result is a device CSR whose
 * rows are primaries, ordered by sky bin.  `on_device` = 1 runs the search
 * on the GPU (HEALPix and plate bins; the result equals the host search's,
 * which runs instead when a plate centre lies within rounding of the
 * annulus edge), 0 runs it on `nthreads` host threads. */
typedef struct {
    uint64_t num_sightline_pairs;        /* considered (search dependent) */
    uint64_t num_sightline_pairs_used;   /* XiEstimator.cc:148 */
    uint64_t num_pixel_pairs;            /* XiEstimator.cc:149 */
    uint64_t num_pixels;                 /* XiEstimator.cc:132 */
    int64_t n_rows;
    int64_t n_skybins;
} tos_pair_stats;

int tos_pairs_build(tos_ctx *ctx, const tos_catalog *cat, const tos_skybins *sky,
                    double max_ang, int nthreads, int on_device, tos_pairs **out);
/* A caller-enumerated candidate list (for SkyBins subclasses the library
 * does not know): rows are primaries, `row_skybin` is the sky-bin key of the
 * row (subsample id), `cos_sep` the caller's angularSeparation value. */
typedef struct {
    int64_t n_rows;
    const int32_t *row_los;
    const int32_t *row_skybin;
    const int64_t *row_ptr;          /* n_rows + 1 */
    const int32_t *other;
    const double *cos_sep;
} tos_pair_csr;
int tos_pairs_upload(tos_ctx *ctx, const tos_catalog *cat, const tos_pair_csr *csr,
                     tos_pairs **out);
int tos_pairs_free(tos_ctx *ctx, tos_pairs *p);
int tos_pairs_stats(const tos_pairs *p, tos_pair_stats *out);
/* Copy the CSR back (arrays sized from tos_pairs_stats; any may be NULL). */
int tos_pairs_download(tos_ctx *ctx, const tos_pairs *p, int32_t *row_los,
                       int32_t *row_skybin, int64_t *row_ptr, int32_t *other,
                       double *cos_sep);

/* ---------------------------------------------------------------------- */
/* The estimator                                                           */
/* ---------------------------------------------------------------------- */

enum tos_precision {
    TOS_FP64 = 0,    /* the reference arithmetic, bin-exact, FP64 sums */
    TOS_FP32C = 1    /* FP32 separations with an FP64 guard-band re-check */
};
enum tos_accum {
    TOS_ACC_CORE = 0, /* num_pairs, sum w1w2d1d2 (didj), sum w1w2 (wgt) */
    TOS_ACC_ALL = 1   /* + di, dj, wi, wj, z: every XiBin field (types.h:70-71) */
};

typedef struct {
    int32_t precision;        /* enum tos_precision */
    int32_t accum;            /* enum tos_accum */
    int32_t per_skybin;       /* 1: also return per-sky-bin subsamples */
    int32_t block_size;       /* 0 = tuned default; else 64..1024, multiple of 64 */
    int32_t blocks_per_cu;    /* 0 = tuned default */
    int32_t flags;            /* TOS_XI_* bits below (0 = all optimisations on) */
    int32_t wg_per_slot;      /* workgroups per resident workgroup slot (0 = default) */
    int32_t reserved;
} tos_xi_opts;

/* tos_xi_opts.flags: switch off individual work-saving schemes (results are
 * identical either way; used for A/B timing). */
enum {
    TOS_XI_NO_WINDOW = 1,     /* evaluate partner pixels outside the r_par window */
    TOS_XI_NO_TRANSPOSE = 2,  /* keep short primary remainders on the lanes */
    TOS_XI_NO_INTERLEAVE = 4, /* consecutive pixels on consecutive lanes */
    TOS_XI_NO_ZCONST = 8      /* bin z per pixel pair even where a whole
                                 sightline pair provably shares one z bin */
};

/* Raw (unfinalized) per-bin sums, the XiBin fields of types.h:70-71.
 * Arrays have getNBinsTotal() entries; di..z are only written with
 * TOS_ACC_ALL (may be NULL otherwise).  Per-sky-bin arrays, when requested,
 * are [n_skybins][nbins] in ascending sky-bin key order. */
typedef struct {
    uint64_t *num_pairs;
    double *didj, *wgt;
    double *di, *dj, *wi, *wj, *z;
    /* optional subsamples (XiEstimator.h:78, save_subsamples :363-383) */
    int32_t *skybin_keys;            /* n_skybins */
    uint64_t *sub_num_pairs;
    double *sub_didj, *sub_wgt;
    /* written by the call */
    uint64_t num_pixel_pairs_used;   /* XiEstimator.cc:151 */
    uint64_t num_guard_rechecks;     /* FP32C pairs re-binned in FP64 */
    uint64_t num_pairs_evaluated;    /* pixel pairs the kernel evaluated (pairs
                                        outside the r_par window are rejected
                                        without evaluation) */
    double kernel_ms;                /* device time of the pair kernel */
} tos_xi_result;

/* XiEstimator::run's pair loop and merge (XiEstimator.cc:31-59,
 * 121-305) for all candidate pairs of `pairs` on the GPU. */
int tos_xi_run(tos_ctx *ctx, const tos_catalog *cat, const tos_pairs *pairs,
               const tos_grid *grid, const tos_xi_opts *opts, tos_xi_result *out);

/* No counterpart in the reference: tos_xi_run (CORE sums) for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation.  It allocates nothing:
 * its workgroup ranges, partial histograms and reduction buffers live in a
 * static device array of 256 MiB per device, held until the process ends;
 * grids too large for it are refused with TOS_ERR_INVALID.  Inputs are a
 * resident catalogue and pair list on which tos_xi_run has run once
 * (building the pair list's work items; for an Observing grid, with this
 * grid); outputs are caller-owned DEVICE arrays of getNBinsTotal() entries.
 * The work is enqueued on `stream` (NULL: the context's stream) and not
 * waited for.  All calls on one device share the static array: calls that
 * overlap on one device race. */
int tos_xi_run_noalloc(tos_ctx *ctx, const tos_catalog *cat, const tos_pairs *pairs,
                       const tos_grid *grid, int32_t precision, uint64_t *num_pairs,
                       double *didj, double *wgt, void *stream);

/* ---------------------------------------------------------------------- */
/* Batched jobs: one persistent launch over many (catalog, pairs) jobs      */
/* ---------------------------------------------------------------------- */

typedef struct {
    const tos_catalog *cat;
    const tos_pairs *pairs;
    tos_grid grid;
    int32_t precision;
    int32_t accum;                   /* TOS_ACC_CORE only in batches */
} tos_batch_job;

/* Runs n jobs through one persistent work-queue kernel (replaces calling
 * XiEstimator::run once per mock).  use_graph = 1 captures the launch in a
 * hipGraph and replays it.  Results: out[k] for job k. */
int tos_batch_run(tos_ctx *ctx, int32_t n, const tos_batch_job *jobs, int use_graph,
                  tos_xi_result *out, double *batch_ms);

/* Times the pair kernel of one job over a small set of launch shapes and
 * keeps the fastest for this context and kernel variant: workgroup size
 * 256/512/1024 x workgroups per CU 1/2/4/8 (as the LDS allows), then
 * workgroups per resident slot 4/8/16/64, then the work-item size
 * (partner pixels per item, 1024-16384; the pair list's items are rebuilt
 * for each).  Later tos_xi_run (block_size = 0), tos_xi_run_per_skybin and
 * tos_batch_run calls use it (the server keeps the default shape).  No
 * counterpart in the reference, whose CUDA launch shape is fixed
 * (gpu/brute.cu:121-149). */
int tos_xi_autotune(tos_ctx *ctx, const tos_catalog *cat, const tos_pairs *pairs,
                    const tos_grid *grid, int32_t precision, int32_t *best_block,
                    int32_t *best_blocks_per_cu, double *best_ms);

/* The shape tos_xi_autotune kept for this grid's kernel variant and
 * precision: shape[0] workgroup size, [1] workgroups per CU, [2] workgroups
 * per resident slot, [3] partner pixels per work item; 0 where the default
 * applies (and all 0 before any tuning).  No counterpart in the reference. */
int tos_xi_tuned_shape(const tos_ctx *ctx, const tos_grid *grid, int32_t precision, int32_t shape[4]);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching of xi jobs                                 */
/* ---------------------------------------------------------------------- */

typedef struct tos_server tos_server;
typedef uint64_t tos_job_id;
typedef struct {
    int64_t hbm_budget_bytes;        /* admission limit for resident buffers */
    int32_t max_inflight;            /* jobs on the device at once */
    int32_t nthreads;                /* host threads for candidate search */
    int32_t search_on_device;        /* 1: candidate search on the GPU (HEALPix,
                                        plates), pair count known before
                                        admission; 0: host search */
    int32_t reserved;
} tos_server_opts;
typedef struct {
    tos_sightlines sightlines;       /* copied at submit */
    tos_skybins sky;
    tos_grid grid;
    double scale;                    /* transverse comoving scale (XiEstimator.h:23) */
    int32_t precision;
} tos_job_desc;
typedef struct {
    double submit_ms, admit_ms, start_ms, done_ms;   /* server clock */
    int64_t footprint_bytes;
    int32_t status;
    uint64_t num_pixel_pairs;        /* candidate pixel pairs (XiEstimator.cc:149) */
} tos_job_timing;

int tos_server_start(tos_ctx *ctx, const tos_server_opts *opts, tos_server **out);
int tos_server_stop(tos_server *srv);
int tos_job_submit(tos_server *srv, const tos_job_desc *job, tos_job_id *id);
int tos_job_poll(tos_server *srv, tos_job_id id);
int tos_job_wait(tos_server *srv, tos_job_id id, double timeout_ms);
int tos_job_cancel(tos_server *srv, tos_job_id id);
/* Copies the finished job's result (CORE sums) and releases its slot.
 * A thread still blocked in tos_job_wait on the same id returns normally. */
int tos_job_result(tos_server *srv, tos_job_id id, tos_xi_result *out,
                   tos_job_timing *timing);
/* The job's timing so far (start_ms stays 0 until its first launch); no
 * counterpart in the reference, which has no job queue. */
int tos_job_timing_get(tos_server *srv, tos_job_id id, tos_job_timing *timing);
int tos_job_footprint(const tos_job_desc *job, int64_t *bytes);

/* ---------------------------------------------------------------------- */
/* 3D brute force (the C0 path)                                            */
/* ---------------------------------------------------------------------- */

enum tos_brute_mode {
    TOS_BRUTE_FP64 = 0,   /* old/BinPair.h:28-47 arithmetic, nbins bins */
    TOS_BRUTE_FP32 = 1    /* gpu/brute.cu:26-55 binning, nbins + 2 slots */
};
/* Replaces bruteGPU (gpu/brute.h:20, gpu/brute.cu:63-224): all pairs i<j of
 * n points (x, y, z, d, w), device resident, no padding of the input.
 * Outputs have nbins (FP64) or nbins + 2 (FP32) entries. */
int tos_brute3d(tos_ctx *ctx, int32_t mode, int64_t n, const double *x, const double *y,
                const double *z, const double *d, const double *w, double min,
                double max, int32_t nbins, double *dsum, double *wsum,
                uint64_t *counts, double *kernel_ms);

/* ---------------------------------------------------------------------- */
/* Covariance of the sky-bin subsamples                                     */
/* ---------------------------------------------------------------------- */

/* Replaces the do_cov block of XiEstimator::run (XiEstimator.cc:60-97,
 * likely::CovarianceAccumulator over the skybin_xis_ samples).  didj and wgt
 * are [n_samples][nbins] (the sub_didj / sub_wgt of a per_skybin run).
 * Sample k: x_kb = didj_kb / wgt_kb where wgt_kb > 0 (else 0), weight
 * w_k = sum of its wgt_kb > 0.  Outputs: mean[nbins] = sum_k w_k x_k / W,
 * cov[nbins][nbins] = sum_k w_k (x_k - mean)(x_k - mean)^T / W, *weight_sum
 * = W.  mean may be NULL.  FP64 on the device (matrix cores). */
int tos_subsample_covariance(tos_ctx *ctx, int64_t n_samples, int32_t nbins,
                             const double *didj, const double *wgt, double *cov,
                             double *mean, double *weight_sum);

/* The covariance diagnostics XiEstimator::run prints
 * (XiEstimator.cc:82-87, likely::CovarianceMatrix::isPositiveDefinite and
 * getLogDeterminant): *is_pos_def = 1 when every pivot of the FP64
 * Cholesky factorisation of the n x n row-major matrix `cov` (its lower
 * triangle is read) is > 0, and then *log_det = log|C| (NaN otherwise).
 * Host only: no context, no device. */
int tos_covariance_check(int32_t n, const double *cov, int32_t *is_pos_def, double *log_det);

/* ---------------------------------------------------------------------- */
/* Cosmology used by the h5healxi distance transform (h5healxi.cc:196-239) */
/* ---------------------------------------------------------------------- */

/* cosmo::LambdaCdmUniverse(OmegaLambda, OmegaMatter): line-of-sight
 * comoving distance and transverse comoving scale in Mpc/h. */
int tos_cosmo_distances(double omega_lambda, double omega_matter, int64_t n,
                        const double *z, double *los_comoving, double *transverse);

/* The per-pixel distance transform of h5healxi.cc:196-216 on the device:
 * distance[i] = (float) D_C(10^(loglam[i] - log10(1216)) - 1) with the
 * cosmology above (omega_matter = 0: flat, h5healxi.cc:124), for n pixels
 * (host arrays; the FP64 arithmetic is the same as tos_cosmo_distances'). */
int tos_pixel_distances(tos_ctx *ctx, double omega_lambda, double omega_matter, int64_t n,
                        const float *loglam, float *distance);

/* ---------------------------------------------------------------------- */
/* Device arithmetic self-test (the FP64 contract relies on these)          */
/* ---------------------------------------------------------------------- */

/* Evaluates on the device: out_sqrt[i] = the pair kernels' FP64 sqrt of
 * a[i] (correctly rounded for a[i] = +-0, +inf, NaN, a[i] < 0 and
 * a[i] >= 2^-767), out_div[i] = a[i]/b[i], out_cvt[i] = x86
 * static_cast<unsigned>(a[i]) emulation. */
int tos_selftest_f64(tos_ctx *ctx, int64_t n, const double *a, const double *b,
                     double *out_sqrt, double *out_div, uint32_t *out_cvt);

#ifdef __cplusplus
}
#endif
#endif /* TOS_AMD_H */
