// tos_internal.h -- shared declarations between the host code, the C-ABI
// layer and the HIP kernels of the gfx950 pair-counting estimator.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include <hip/hip_runtime.h>

#include "tos_amd.h"

namespace tos {

// One sightline on the device: pixel offset/count and the loglam range
// (used to prove that a whole sightline pair falls in one z bin).
struct LosInfo {
    int32_t off;
    int32_t n;
    float lmin;
    float lmax;
};

// Unit of work of one wavefront: primary pixels of sightline `los` against
// partner sightlines pair_begin..pair_end-1.  Normal items put the 64
// primary pixels [p1_begin, p1_begin+64) on the lanes and loop over the
// partner pixels; transposed items (p1_begin has kTransposed set) put 64
// partner pixels on the lanes and loop over the short primary remainder
// [p1_begin, n_los).
struct XiItem {
    int32_t los;
    int32_t p1_begin;
    int32_t pair_begin;
    int32_t pair_end;
};
constexpr int32_t kTransposed = 0x40000000;
// largest partner-pixel run of one work item
constexpr int32_t kItemStepCap = 4096;
// remainders of fewer primary pixels than this run transposed
constexpr int32_t kTransposeBelow = 48;

constexpr uint32_t kReject = 0xFFFFFFFFu;

// Grid constants for the exact (reference) arithmetic.
struct GridExact {
    double min[3];
    double inv[3];        // 1/width for axes 0,1; 1/NBins for axis 2 (XiEstimator.cc:232)
    uint32_t n[3];
    uint32_t ntot;
    // LDS histogram layout: one row of `row` slots per axis-0 bin, with
    // row = n1*n2 + pad odd, so that neighbouring axis-0 bins (the bins that
    // the lanes of a wavefront hit together) fall in different LDS banks.
    uint32_t rowlen;      // n1 * n2
    uint32_t pad;
    uint32_t nslots;      // n0 * (rowlen + pad)
    uint32_t row8;        // 8 (rowlen + pad): byte stride of an axis-0 row of FP64 sums
    uint32_t nz8;         // 8 n2
};

// Constants of the FP32 fast path with its guard band.  The kernel forms
// u_k = x * inv[k] + offg[k] = t_k + g_k (t_k the bin coordinate, g_k the
// guard) and re-checks a pair when fract(u_k) < guard2[k] = 2 g_k, i.e. when
// t_k lies within g_k of a bin edge.  Outside that band trunc(u_k) is the bin
// of t_k and u_k passes the same range tests as t_k.
struct GridFast {
    float inv[2];
    float offg[2];        // -min * inv + g (axis 0: -min * inv when exact0)
    float nf[2];          // bin counts as float
    float guard2[2];      // 2 g_k (>= 1: always re-check)
    float small;          // re-check when r_perp^2 (Cartesian) / r^2 (Polar) is below this
    uint32_t exact0;      // 1: axis-0 t is exact in FP32 (no guard on axis 0)
    uint32_t nonneg;      // bit k: t_k >= 0 for every pair (skip the > -1 test)
    uint32_t small_safe;  // 1 (Cartesian): r_perp^2 < small implies bin 0 in both arithmetics
    uint32_t zero_min;    // bit k: axis k starts at exactly 0, so a zero coordinate is bin 0 exactly
};

struct XiKernelArgs {
    const float4 *pix;            // {distance, loglam, weight, weight*value}
    const LosInfo *los;
    const int32_t *other;         // partner sightline of each candidate pair
    const double *csep;           // cos(angular separation) of each pair
    const uint32_t *theta_bin;    // Observing grid: theta bin of each pair
    const XiItem *items;
    const int32_t *wg_item_ptr;   // [n_wg + 1] item range of each workgroup
    uint32_t *part_cnt;           // [n_wg][nbins]
    double *part_sum;             // [n_wg][nbins][nsum]
    unsigned long long *n_recheck;   // [0] guard re-checks, [1] evaluated pixel pairs
    GridExact ge;
    GridFast gf;
    int32_t nbins;
    // r_par window: a pixel pair whose axis-0 key (distance, or loglam for
    // the Observing grid) differs by >= win_half is rejected by the
    // reference arithmetic, so sorted partner runs are clipped to
    // (key_min - win_half, key_max + win_half).  window = 0: off.
    int32_t window;
    double win_half;
    // distance window: with window = 1 and perp_cut > 0, partner pixels
    // with d2 >= perp_cut / (2 (1 - c) d1_min) are rejected by the reference
    // (grid_perp_cut) and skipped as well
    double perp_cut;
    // 1: interleaved lane <-> pixel order in 64-pixel chunks (chunk_pixel)
    int32_t interleave;
    // 1: sightline pairs whose extreme loglam sums share one z bin take
    // that bin as a constant (pair_z_const); 0: z binned per pixel pair
    int32_t zconst;
    // The slots a dynamic-LDS launch histograms: byte offsets
    // [band_lo8, band_lo8 + band_len8) of the slot array, flat bins
    // band_bin0 + [0, nbins).  The whole grid unless `banded`: then a band of
    // axis-0 rows of a grid too large for one LDS histogram (always the
    // dynamic-LDS kernel variant).
    uint32_t band_lo8;
    uint32_t band_len8;
    int32_t band_bin0;
    int32_t banded;
};

// One workgroup's task in a batched launch: items [begin, end) of job
// `job`, written to partial row `row` of that job.
struct XiWgTask {
    int32_t job;
    int32_t begin;
    int32_t end;
    int32_t row;
};

// A batched launch: per-job arguments (each with its own part_cnt /
// part_sum / n_recheck) and one task per workgroup.  Every job of one
// launch uses the same kernel variant (coordinates, precision, layout).
struct XiBatchArgs {
    const XiKernelArgs *jobs;
    const XiWgTask *wg;
};

// Per-job input/output of the batched reduction.
struct XiReduceJob {
    const uint32_t *part_cnt;     // [n_wg][nbins]
    const double *part_sum;       // [n_wg][nbins][nsum]
    int32_t n_wg;
    int32_t nbins;
    int32_t nsum;
    int32_t pad;
    uint64_t *out_cnt;            // [nbins]
    double *out_sum;              // [nbins][nsum]
};

// Launchers (tos_kernels.hip). All return hipError_t as int.
int launch_xi(int coord, int precision, int accum, const XiKernelArgs &a, int n_wg,
              int block, size_t lds_bytes, void *stream);
// Kernel variant of a job: `fixed` LDS layout and `simple` grid (see the
// kernel file); jobs of one batched launch must agree on both.
void xi_variant(int coord, const GridExact &ge, const GridFast &gf, bool &fixed, bool &simple);
// n_wg == 0 sets the variant's dynamic LDS limit to lds_bytes and launches
// nothing; call it once per variant before launching (or capturing) it.
int launch_xi_batch(int coord, int precision, bool fixed, bool simple, const XiBatchArgs &b,
                    int n_wg, int block, size_t lds_bytes, void *stream);
int launch_reduce_batch(const XiReduceJob *jobs_dev, int n_jobs, int64_t max_fields, void *stream);
// Device candidate search (tos_search.hip): one wavefront per primary.
struct SearchArgs {
    const double4 *trig;          // per sightline: {sin dec, cos dec, sin ra, cos ra}
    const LosInfo *los;           // pixel counts
    const int32_t *prim;          // [n_prim] primary sightline of each row candidate, row order
    const int32_t *prim_cell;     // [n_prim] its occupied-bin index
    const int32_t *cell_ptr;      // [n_bins + 1] members of each occupied bin
    const int32_t *members;
    const int32_t *nbr_ptr;       // [n_bins + 1] neighbour bins of each occupied bin
    const int32_t *nbr;
    double cos_max;
    int64_t n_prim;
    // count pass
    int32_t *count;               // [n_prim] kept partners (written by the count pass, read by the fill pass)
    uint64_t *considered;         // [n_prim] candidates j > i examined
    uint64_t *pixel_pairs;        // [n_prim] n_i * sum of the kept partners' n_j
    // fill pass
    const int64_t *offset;        // [n_prim] first output slot of each primary
    int32_t *other;
    double *csep;
    // plate sky bins (nbr_ptr / nbr unused): plates in ascending id order
    int32_t plates;               // 1: plate neighbour test instead of nbr lists
    int32_t n_plate;
    const double4 *plate_trig;    // [n_plate]
    const int32_t *plate_bin;     // [n_plate] occupied-bin index of the plate, -1 if none
    // |acos(c) - 1.5 deg| < max_ang as c in (c_up, c_lo), open; has_up /
    // has_lo = 0 when that side is vacuous; a c within `margin` of a bound is
    // undecidable against the host acos and raises *undecided
    double c_up, c_lo, margin;
    int32_t has_up, has_lo;
    int32_t *undecided;           // device flag, 0 on entry
};
int launch_search(const SearchArgs &a, bool fill, void *stream);

// Work items on the device (the build_items rules, one thread per row):
// the count pass writes the item count of every row, the fill pass the
// items and their costs from row_item_ptr on.
struct ItemArgs {
    const int32_t *row_los;
    const int64_t *row_ptr;       // [n_rows + 1]
    const int32_t *other;
    const LosInfo *los;
    int64_t n_rows;
    int32_t step_cap;
    int32_t transpose_below;
    int32_t *count;               // count pass: [n_rows]
    const int32_t *row_item_ptr;  // fill pass: [n_rows + 1]
    XiItem *items;
    uint64_t *item_cost;
};
int launch_items(const ItemArgs &a, bool fill, void *stream);

// Zero n 64-bit words with a kernel (safe inside a captured graph).
int launch_zero_u64(unsigned long long *p, int64_t n, void *stream);
int launch_reduce(const uint32_t *part_cnt, const double *part_sum, int n_wg, int nbins,
                  int nsum, uint64_t *out_cnt, double *out_sum, uint32_t *scratch_cnt,
                  double *scratch_sum, int n_seg, void *stream);
// 3D brute force (tos_brute.hip).  pts/w: packed points {x, y, z, delta}
// and weights; tiles: {i_block_begin, j_begin} pairs, each tile the 64
// points of the block against points [j_begin, min(n, j_begin + chunk)).
struct BruteArgs {
    const double4 *pts;           // {x, y, z, delta} per point
    const double *w;              // weight per point
    int64_t n;
    double mn, mx;
    int32_t nbins;
    int32_t copies;               // histogram copies per workgroup (power of two, <= 64)
    const int32_t *tiles;
    int32_t n_tiles;
    int32_t chunk;
    double inv_width;             // 1 / FP64 bin width when it is a power of two
    float inv_spacing;            // 1 / FP32 spacing when it is a power of two
    uint32_t *part_cnt;           // [n_wg][nslot]
    double *part_sum;             // [n_wg][nslot][2] {dsum, wsum}
};
size_t brute3d_lds_bytes(int mode, int32_t nbins, int copies);
// Weighted covariance of k_n samples of n bins (tos_cov.hip); x, a: k_n*n
// scratch, w: k_n, mu: n, wsum: 1, cov: n*n, all device memory.
int launch_covariance(const double *didj, const double *wgt, int64_t k_n, int32_t n, double *x, double *a,
                      double *w, double *mu, double *wsum, double *cov, void *stream);
int launch_brute3d(int mode, bool pow2, const BruteArgs &a, int n_wg, void *stream);
int launch_selftest(const double *a, const double *b, double *s, double *q, uint32_t *c,
                    int64_t n, void *stream);
size_t xi_lds_bytes(int accum, int nslots);
int xi_nsum(int accum);
int xi_max_lds(int device);

// ---- host algorithms (tos_host.cpp) ----

struct HostPairs {
    std::vector<int32_t> row_los;
    std::vector<int32_t> row_skybin;
    std::vector<int64_t> row_ptr;   // n_rows + 1
    std::vector<int32_t> other;
    std::vector<double> cos_sep;
    std::vector<int32_t> skybin_keys;   // every occupied sky bin, ascending
    tos_pair_stats stats{};
};

// HEALPix RING ang2pix (theta colatitude, phi longitude).
int64_t healpix_ang2pix_ring(int64_t nside, double theta, double phi);
void healpix_pix2ang_ring(int64_t nside, int64_t pix, double *theta, double *phi);
int64_t healpix_radec2pix(int64_t nside, double ra, double dec);

// SkyObject trigonometry (types.h:20-26), FP64 with the host libm.
struct Trig {
    double sd, cd, sr, cr;
};
Trig make_trig(double ra, double dec);

// The sky binning of a catalogue: occupied bins (first-insertion order,
// SkyBins.h:59-61), their members in sightline order, the bins in
// ascending key order, and, for HEALPix, every bin's neighbour bins (a
// superset of the cells the reference's inclusive disc query returns, in
// ascending key order).  Plates keep their centres for the per-sightline
// annulus test.
struct SkyIndex {
    std::vector<Trig> trig;
    std::vector<int32_t> occupied;
    std::unordered_map<int32_t, int32_t> slot;
    std::vector<std::vector<int32_t>> members;
    std::vector<int32_t> by_key;
    std::vector<std::vector<int32_t>> nbr;
    std::vector<Trig> plate_trig;
    std::vector<int32_t> plate_order;
};
void sky_index(const tos_sightlines &s, const tos_skybins &sky, double max_ang, SkyIndex &ix);

// Candidate enumeration over sky bins (XiEstimator.cc:121-159).
int build_pairs_host(const tos_sightlines &s, const tos_skybins &sky, double max_ang,
                     int nthreads, HostPairs &out, std::string &err);

// Balances work items (costs item_cost) over n_wg workgroups: the
// [n_wg + 1] item pointer of contiguous ranges of equal cost.
void balance_items(const std::vector<uint64_t> &item_cost, int n_wg,
                   std::vector<int32_t> &wg_item_ptr);

// Exact-arithmetic grid constants and FP32 guard constants.
void grid_exact_consts(const tos_grid &g, GridExact &ge);
void grid_fast_consts(const tos_grid &g, double dmin, double dmax, GridFast &gf);
// Smallest axis-0 key difference that the reference arithmetic always
// rejects (see XiKernelArgs::win_half); < 0 when no such bound is usable.
// kmax bounds |key| (distance, or loglam for the Observing grid).
double grid_window_half(const tos_grid &g, double dmax, double kmax);
// A pixel pair with 2 d1 d2 (1 - c) >= this is rejected by the reference
// (its r_perp, Cartesian, or r, Polar, lies beyond the last bin); < 0 when
// the grid has no such bound (Observing).
double grid_perp_cut(const tos_grid &g, double dmin, double dmax);

// x86-64 static_cast<unsigned>(double) on the host.
uint32_t cvt_x86(double x);

// Flat LCDM (with curvature) distances in Mpc/h.
double cosmo_los_comoving(double omega_lambda, double omega_matter, double z);
double cosmo_transverse(double omega_lambda, double omega_matter, double z);

}  // namespace tos
