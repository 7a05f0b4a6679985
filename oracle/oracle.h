/*
 * oracle.h -- CPU restatement of turbo-octo-spice's pair-counting xi estimator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product never links or calls it.
 *
 * Every function restates the reference's arithmetic in source order with
 * no FMA contraction (built with -O2 -ffp-contract=off, no fast-math):
 *   - XiBin accumulate/finalize/merge       turbooctospice/types.h:69-121
 *   - SkyObject::angularSeparation          turbooctospice/types.h:20-39
 *   - XiEstimator ctor / run / task / merge turbooctospice/XiEstimator.cc:19-159, 298-305
 *   - accumulate_pixel_pairs (3 branches)   turbooctospice/XiEstimator.cc:161-296
 *   - maxAngularScale                       CartesianGrid.h:46, PolarGrid.h:46, QuasarGrid.h:45
 *   - PlateBins neighbour rule              turbooctospice/PlateBins.h:74-85
 *   - HEALPix RING ang2pix / pix2ang        HealpixBins.h:58-72 (healpix_cxx 3.30, Gorski+2005)
 *   - 3D brute force, FP64                  turbooctospice/old/BinPair.h:28-47,
 *                                           old/BrutePairSearch.h:31-50 (pair order)
 *   - 3D brute force, FP32 binning          gpu/brute.cu:26-55, :173-186, :206-210
 *
 * Pinning status of each function is listed in the header comment of the
 * file that defines it.
 */
#ifndef TOS_ORACLE_H
#define TOS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* XiBin (types.h:69-71): 7 double sums + a long pair count. */
typedef struct {
    double didj, di, dj, wgt, wi, wj, z;
    int64_t num_pairs;
} ora_xibin;

/* Sightlines in structure-of-arrays form.  LOS k owns pixels
 * [offsets[k], offsets[k+1]).  ra/dec are radians (types.h:18-19). */
typedef struct {
    int64_t n_los;
    const int64_t *offsets;
    const double *ra, *dec;
    const float *value, *loglam, *weight, *distance;
} ora_sightlines;

/* Coordinate type, same order as XiEstimator::BinningCoordinateType
 * (XiEstimator.h:19-21). */
enum { ORA_POLAR = 0, ORA_CARTESIAN = 1, ORA_OBSERVING = 2 };

/* A 3-axis uniform grid: per axis the low edge of bin 0, the bin width,
 * the high edge of the last bin and the bin count
 * (AbsTwoPointGrid.h:55-58, 71-85). */
typedef struct {
    int type;
    int32_t nbins[3];
    double min[3], width[3], max[3];
} ora_grid;

/* Sky binning used to enumerate candidate sightline pairs. */
enum { ORA_SKY_HEALPIX = 0, ORA_SKY_PLATES = 1 };
typedef struct {
    int kind;
    const int32_t *bin_of_los;      /* sky-bin key of each LOS (HEALPix cell or plate id) */
    /* PlateBins only: the plate list (PlateBins.h:40-45) */
    int64_t n_plates;
    const int32_t *plate_id;
    const double *plate_ra, *plate_dec;   /* radians */
} ora_skybins;

typedef struct {
    uint64_t num_sightline_pairs_used;   /* XiEstimator.cc:148 */
    uint64_t num_pixel_pairs;            /* XiEstimator.cc:149 */
    uint64_t num_pixel_pairs_used;       /* XiEstimator.cc:151 */
    uint64_t num_pixels;                 /* XiEstimator.cc:132 */
    uint64_t num_sightline_pairs_considered;  /* depends on the neighbour search */
} ora_stats;

/* x86-64 static_cast<unsigned>(double): cvttsd2si to 64 bits, low 32 bits. */
uint32_t ora_cvt_u32(double x);

/* XiBin::accumulate_pair (types.h:77-87). */
void ora_accumulate_pair(ora_xibin *b, float v1, float l1, float w1,
                         float v2, float l2, float w2);
/* XiBin::finalize (types.h:96-101). */
void ora_finalize(ora_xibin *b);

/* SkyObject::angularSeparation (types.h:37-39) from ra/dec in radians. */
double ora_cos_separation(double ra1, double dec1, double ra2, double dec2);

/* AbsTwoPointGrid::maxAngularScale for the grid's coordinate type. */
double ora_max_angular_scale(const ora_grid *g, double scale);

/* accumulate_pixel_pairs for one sightline pair (XiEstimator.cc:161-296);
 * `xi` holds getNBinsTotal() bins; returns num_pixel_pairs_used. */
uint64_t ora_accumulate_los_pair(const ora_sightlines *s, const ora_grid *g,
                                 int64_t primary, int64_t other,
                                 double cos_separation, ora_xibin *xi);

/* XiEstimator::run (XiEstimator.cc:31-59, 121-159) without covariance.
 *   xi_final   : getNBinsTotal() bins after merge + finalize (may be NULL)
 *   xi_raw     : same, merged but not finalized (may be NULL)
 *   per_skybin : [n_skybins][nbins] unfinalized subsample sums, ordered by
 *                ascending sky-bin key (std::map order) (may be NULL);
 *                skybin_keys receives the keys (may be NULL)
 *   task_stride/task_phase: run only occupied-bin tasks t with
 *                t % stride == phase (1/0 = all); used to time bounded samples.
 * Returns the number of occupied sky bins, or -1 on bad input. */
int64_t ora_xi_run(const ora_sightlines *s, const ora_grid *g, double scale,
                   const ora_skybins *sky, int nthreads,
                   int64_t task_stride, int64_t task_phase,
                   ora_xibin *xi_final, ora_xibin *xi_raw,
                   ora_xibin *per_skybin, int32_t *skybin_keys,
                   ora_stats *stats);

/* HEALPix RING scheme, nside = 2^order. theta = colatitude, phi = longitude. */
int64_t ora_ang2pix_ring(int64_t nside, double theta, double phi);
void ora_pix2ang_ring(int64_t nside, int64_t pix, double *theta, double *phi);
/* HealpixBins::ang2pix with radec2pnt (HealpixBins.h:58-72). */
int64_t ora_radec2pix(int64_t nside, double ra, double dec);

/* 3D isotropic brute force, FP64 (old/BinPair.h:28-47): every pair i<j.
 * dsum/wsum/counts have nbins entries, zeroed by the callee. */
void ora_brute3d_f64(const double *x, const double *y, const double *z,
                     const double *d, const double *w, int64_t n,
                     double min, double max, int32_t nbins,
                     double *dsum, double *wsum, int64_t *counts);

/* 3D brute force with gpu/brute.cu's FP32 separation and bin index
 * (brute.cu:26-55).  hist has nbins+2 slots (underflow, bins, overflow);
 * sums are accumulated in double so the result is order independent to
 * ~1e-15; counts are exact. */
void ora_brute3d_f32(const float *x, const float *y, const float *z,
                     const float *d, const float *w, int64_t n,
                     float min, float max, int32_t nbins,
                     double *dsum, double *wsum, int64_t *counts);

#ifdef __cplusplus
}
#endif
#endif
