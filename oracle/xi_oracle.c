/*
 * xi_oracle.c -- CPU restatement of XiEstimator (turbooctospice/XiEstimator.cc).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -O2 -ffp-contract=off
 * and without fast-math so every float/double operation happens once, in
 * the reference's source order.
 *
 * Pinning (tests/test_oracle_*.py):
 *   ora_cvt_u32            x86 cvttsd2si probes: NaN, -0.5, -1.5, -4.9, 2^32+5
 *   ora_accumulate_pair    hand-built pixel pairs, closed-form sums
 *   ora_cos_separation     spherical law of cosines vs 3D unit-vector dot product
 *   Cartesian/Polar/Observing branches
 *                          closed-form geometry: 3D positions in long double,
 *                          bins away from edges; E8 NaN path and E9 negative
 *                          casts on hand-made pairs; inv_z = 1/NBins quirk
 *   ora_xi_run             candidate search: an O(N^2) numpy enumeration of
 *                          every i < j with c > cos_max (partners placed at
 *                          cos_max - 1 ulp, cos_max, + 1 ulp; a copy of a
 *                          sightline at c ~ 1) gives the same pair statistics
 *                          and per-sky-bin histograms at HEALPix orders 0-6
 *                          (tests/test_oracle_search_pin.py);
 *                          sum(counts) == num_pixel_pairs_used; LOS relabel
 *                          invariance; PlateBins annulus rule on a hand case
 * Pinned only by the checks above (no reference golden vectors exist):
 *   summation order inside a sky bin after the HEALPix query (ascending
 *   cell id assumed for healpix_cxx's query_disc_inclusive) -- affects FP64
 *   sums at the 1e-16 relative level only.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

uint32_t ora_cvt_u32(double x)
{
    /* cvttsd2si with a 64-bit destination: NaN and values whose truncation
     * does not fit in int64 give 0x8000000000000000, whose low half is 0. */
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)x;
}

void ora_accumulate_pair(ora_xibin *b, float v1, float l1, float w1,
                         float v2, float l2, float w2)
{
    /* types.h:78-86 */
    float wdi = w1 * v1;
    float wdj = w2 * v2;
    float didj = wdi * wdj;
    float ww = w1 * w2;
    float lsum = l1 + l2;
    b->didj += (double)didj;
    b->di += (double)wdi;
    b->dj += (double)wdj;
    b->wgt += (double)ww;
    b->wi += (double)w1;
    b->wj += (double)w2;
    b->z += 0.5 * (double)lsum;
    b->num_pairs += 1;
}

void ora_finalize(ora_xibin *b)
{
    /* types.h:96-101 */
    if (b->wgt > 0) b->didj /= b->wgt; else b->didj = 0;
    if (b->wi > 0) b->di /= b->wi; else b->di = 0;
    if (b->wj > 0) b->dj /= b->wj; else b->dj = 0;
    if (b->num_pairs > 0) b->z /= (double)b->num_pairs; else b->z = 0;
}

static void xibin_add(ora_xibin *a, const ora_xibin *b)
{
    /* types.h:104-113 */
    a->didj += b->didj;
    a->di += b->di;
    a->dj += b->dj;
    a->wgt += b->wgt;
    a->wi += b->wi;
    a->wj += b->wj;
    a->z += b->z;
    a->num_pairs += b->num_pairs;
}

typedef struct {
    double sin_dec, cos_dec, sin_ra, cos_ra;
} sky_trig;

static sky_trig make_trig(double ra, double dec)
{
    /* types.h:20-26 */
    sky_trig t;
    t.sin_dec = sin(dec);
    t.cos_dec = cos(dec);
    t.sin_ra = sin(ra);
    t.cos_ra = cos(ra);
    return t;
}

static double trig_cos_sep(const sky_trig *a, const sky_trig *b)
{
    /* types.h:37-39, evaluated left to right */
    double t1 = a->sin_dec * b->sin_dec;
    double t2 = a->cos_dec * b->cos_dec;
    double t3 = a->sin_ra * b->sin_ra;
    double t4 = a->cos_ra * b->cos_ra;
    return t1 + t2 * (t3 + t4);
}

double ora_cos_separation(double ra1, double dec1, double ra2, double dec2)
{
    sky_trig a = make_trig(ra1, dec1), b = make_trig(ra2, dec2);
    return trig_cos_sep(&a, &b);
}

double ora_max_angular_scale(const ora_grid *g, double scale)
{
    switch (g->type) {
    case ORA_CARTESIAN: /* CartesianGrid.h:46 */
        return g->max[1] / scale;
    case ORA_POLAR:     /* PolarGrid.h:46 */
        return g->max[0] * sqrt(1.0 - g->min[1] * g->min[1]) / scale;
    default:            /* QuasarGrid.h:45: the theta axis is used as radians */
        return g->max[1];
    }
}

uint64_t ora_accumulate_los_pair(const ora_sightlines *s, const ora_grid *g,
                                 int64_t primary, int64_t other,
                                 double cos_separation, ora_xibin *xi)
{
    const float *dist = s->distance, *ll = s->loglam, *wt = s->weight, *val = s->value;
    const int64_t a0 = s->offsets[primary], a1 = s->offsets[primary + 1];
    const int64_t b0 = s->offsets[other], b1 = s->offsets[other + 1];
    const uint32_t n_total = (uint32_t)g->nbins[0] * (uint32_t)g->nbins[1] * (uint32_t)g->nbins[2];
    uint64_t used = 0;
    const double separation = acos(cos_separation);   /* XiEstimator.cc:166 */

    if (g->type == ORA_POLAR) {
        /* XiEstimator.cc:168-224 */
        const double r_min = g->min[0], one_over_dr = 1.0 / g->width[0];
        const double mu_min = g->min[1], one_over_dmu = 1.0 / g->width[1];
        const double z_min = g->min[2], one_over_dz = 1.0 / (double)g->nbins[2];
        const uint32_t nr = (uint32_t)g->nbins[0], nmu = (uint32_t)g->nbins[1], nz = (uint32_t)g->nbins[2];
        for (int64_t p = a0; p < a1; ++p) {
            const float d1 = dist[p];
            const float d1sq = d1 * d1;
            const double r1_sq = (double)d1sq;
            const float two_d1 = 2.0f * d1;
            const double two_r1_cos12 = (double)two_d1 * cos_separation;
            for (int64_t q = b0; q < b1; ++q) {
                const float d2f = dist[q];
                const double d2 = (double)d2f;
                const double r_sq = r1_sq + (d2 - two_r1_cos12) * d2;
                const double r = sqrt(r_sq);
                uint32_t bin = ora_cvt_u32((r - r_min) * one_over_dr);
                if (bin >= nr) continue;
                if (r == 0) continue;
                const float adiff = fabsf(d1 - d2f);
                const double mu = (double)adiff / r;
                const uint32_t mubin = ora_cvt_u32((mu - mu_min) * one_over_dmu);
                if (mubin >= nmu) continue;
                bin = mubin + bin * nmu;
                const float lsum = ll[p] + ll[q];
                const double zc = 0.5 * (double)lsum;
                const uint32_t zbin = ora_cvt_u32((zc - z_min) * one_over_dz);
                bin = zbin + bin * nz;
                if (bin >= n_total) continue;   /* out-of-bounds write in the reference */
                ++used;
                ora_accumulate_pair(&xi[bin], val[p], ll[p], wt[p], val[q], ll[q], wt[q]);
            }
        }
    } else if (g->type == ORA_CARTESIAN) {
        /* XiEstimator.cc:226-259 */
        const double rpar_min = g->min[0], one_over_drpar = 1.0 / g->width[0];
        const double rperp_min = g->min[1], one_over_drperp = 1.0 / g->width[1];
        const double z_min = g->min[2], one_over_dz = 1.0 / (double)g->nbins[2];
        const uint32_t npar = (uint32_t)g->nbins[0], nperp = (uint32_t)g->nbins[1], nz = (uint32_t)g->nbins[2];
        for (int64_t p = a0; p < a1; ++p) {
            const float d1 = dist[p];
            const float d1sq = d1 * d1;
            const double r1_sq = (double)d1sq;
            const float two_d1 = 2.0f * d1;
            const double two_r1_cos12 = (double)two_d1 * cos_separation;
            for (int64_t q = b0; q < b1; ++q) {
                const float d2f = dist[q];
                const double r_para = (double)fabsf(d1 - d2f);
                uint32_t bin = ora_cvt_u32((r_para - rpar_min) * one_over_drpar);
                if (bin >= npar) continue;
                const double d2 = (double)d2f;
                const double r_sq = r1_sq + (d2 - two_r1_cos12) * d2;
                const double rp2 = fabs(r_sq - r_para * r_para);
                const uint32_t perp = ora_cvt_u32((sqrt(rp2) - rperp_min) * one_over_drperp);
                if (perp >= nperp) continue;
                bin = perp + bin * nperp;
                const float lsum = ll[p] + ll[q];
                const double zc = 0.5 * (double)lsum;
                const uint32_t zbin = ora_cvt_u32((zc - z_min) * one_over_dz);
                bin = zbin + bin * nz;
                if (bin >= n_total) continue;
                ++used;
                ora_accumulate_pair(&xi[bin], val[p], ll[p], wt[p], val[q], ll[q], wt[q]);
            }
        }
    } else {
        /* XiEstimator.cc:261-294 */
        const double dl_min = g->min[0], one_over_ddl = 1.0 / g->width[0];
        const double th_min = g->min[1], one_over_dth = 1.0 / g->width[1];
        const double z_min = g->min[2], one_over_dz = 1.0 / (double)g->nbins[2];
        const uint32_t ndl = (uint32_t)g->nbins[0], nth = (uint32_t)g->nbins[1], nz = (uint32_t)g->nbins[2];
        const uint32_t sep_bin = ora_cvt_u32((separation - th_min) * one_over_dth);
        for (int64_t p = a0; p < a1; ++p) {
            for (int64_t q = b0; q < b1; ++q) {
                const float adl = fabsf(ll[p] - ll[q]);
                uint32_t bin = ora_cvt_u32(((double)adl - dl_min) * one_over_ddl);
                if (bin >= ndl) continue;
                bin = sep_bin + bin * nth;
                const float lsum = ll[p] + ll[q];
                const double zc = 0.5 * (double)lsum;
                const uint32_t zbin = ora_cvt_u32((zc - z_min) * one_over_dz);
                bin = zbin + bin * nz;
                if (bin >= n_total) continue;
                ++used;
                ora_accumulate_pair(&xi[bin], val[p], ll[p], wt[p], val[q], ll[q], wt[q]);
            }
        }
    }
    return used;
}

/* ------------------------------------------------------------------ */
/* Sky bins: keys in first-insertion order, members in LOS order       */
/* (SkyBins.h:18,59-61; HealpixBins.h:31-43; PlateBins.h:56-70).      */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t n_keys;
    int32_t *keys_insertion;   /* occupied bins, first-insertion order */
    int32_t *keys_sorted;      /* same keys ascending (std::map order) */
    int64_t *member_ptr;       /* CSR over keys_sorted */
    int64_t *members;          /* LOS ids, ascending inside a bin */
} bin_index;

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

static int64_t find_key(const bin_index *bi, int32_t key)
{
    int64_t lo = 0, hi = bi->n_keys;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (bi->keys_sorted[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < bi->n_keys && bi->keys_sorted[lo] == key) ? lo : -1;
}

static int build_bin_index(const int32_t *bin_of_los, int64_t n_los, bin_index *bi)
{
    memset(bi, 0, sizeof(*bi));
    int32_t *sorted = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_los > 0 ? n_los : 1));
    if (!sorted) return -1;
    memcpy(sorted, bin_of_los, sizeof(int32_t) * (size_t)n_los);
    qsort(sorted, (size_t)n_los, sizeof(int32_t), cmp_i32);
    int64_t nk = 0;
    for (int64_t i = 0; i < n_los; ++i)
        if (i == 0 || sorted[i] != sorted[i - 1]) sorted[nk++] = sorted[i];
    bi->n_keys = nk;
    bi->keys_sorted = sorted;
    bi->member_ptr = (int64_t *)calloc((size_t)nk + 1, sizeof(int64_t));
    bi->members = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_los > 0 ? n_los : 1));
    bi->keys_insertion = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nk > 0 ? nk : 1));
    char *seen = (char *)calloc((size_t)nk + 1, 1);
    if (!bi->member_ptr || !bi->members || !bi->keys_insertion || !seen) { free(seen); return -1; }
    int64_t ni = 0;
    for (int64_t i = 0; i < n_los; ++i) {
        int64_t k = find_key(bi, bin_of_los[i]);
        bi->member_ptr[k + 1]++;
        if (!seen[k]) { seen[k] = 1; bi->keys_insertion[ni++] = bin_of_los[i]; }
    }
    for (int64_t k = 0; k < nk; ++k) bi->member_ptr[k + 1] += bi->member_ptr[k];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nk > 0 ? nk : 1));
    for (int64_t k = 0; k < nk; ++k) fill[k] = bi->member_ptr[k];
    for (int64_t i = 0; i < n_los; ++i) {
        int64_t k = find_key(bi, bin_of_los[i]);
        bi->members[fill[k]++] = i;
    }
    free(fill);
    free(seen);
    return 0;
}

static void free_bin_index(bin_index *bi)
{
    free(bi->keys_insertion);
    free(bi->keys_sorted);
    free(bi->member_ptr);
    free(bi->members);
}

/* ------------------------------------------------------------------ */
/* The estimator run                                                   */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t j;
    int32_t key;
    double c;
} cand;

static int cmp_cand(const void *a, const void *b)
{
    const cand *x = (const cand *)a, *y = (const cand *)b;
    if (x->key != y->key) return (x->key > y->key) - (x->key < y->key);
    return (x->j > y->j) - (x->j < y->j);
}

typedef struct {
    const ora_sightlines *s;
    const ora_grid *g;
    const ora_skybins *sky;
    const bin_index *bi;
    const sky_trig *trig;
    const sky_trig *plate_trig;
    const int64_t *by_dec;        /* LOS ids sorted by dec */
    const double *dec_sorted;
    double max_ang, cos_max;
    uint32_t nbins;
    int64_t n_tasks;              /* tasks actually run */
    const int64_t *task_key_idx;  /* index into keys_sorted for each task */
    ora_xibin *task_xi;           /* [n_tasks][nbins] */
    ora_stats *task_stats;
    int64_t next;                 /* atomic task counter */
} run_ctx;

static int64_t lower_bound_d(const double *a, int64_t n, double v)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void run_task(run_ctx *rc, int64_t t)
{
    const ora_sightlines *s = rc->s;
    const bin_index *bi = rc->bi;
    ora_xibin *xi = rc->task_xi + (size_t)t * rc->nbins;
    ora_stats st;
    memset(&st, 0, sizeof(st));
    memset(xi, 0, sizeof(ora_xibin) * rc->nbins);
    const int64_t k = rc->task_key_idx[t];
    size_t cap = 1024;
    cand *cs = (cand *)malloc(sizeof(cand) * cap);
    const double deg2rad = atan(1.0) * 4 / 180.0;   /* constants.h:12-13 */

    for (int64_t m = bi->member_ptr[k]; m < bi->member_ptr[k + 1]; ++m) {
        const int64_t i = bi->members[m];
        const int64_t ni = s->offsets[i + 1] - s->offsets[i];
        st.num_pixels += (uint64_t)ni;
        size_t nc = 0;
        if (rc->sky->kind == ORA_SKY_HEALPIX) {
            /* Every j > i with c > cos_max, visited by ascending (cell, j):
             * the inclusive disc query returns a superset of the cells that
             * hold such j, in ascending cell order. The angular distance is at
             * least |dec_i - dec_j|, so a dec window bounds the search. */
            const double margin = 1e-7;
            int64_t lo = lower_bound_d(rc->dec_sorted, s->n_los, s->dec[i] - rc->max_ang - margin);
            int64_t hi = lower_bound_d(rc->dec_sorted, s->n_los, s->dec[i] + rc->max_ang + margin);
            for (int64_t q = lo; q < hi; ++q) {
                const int64_t j = rc->by_dec[q];
                if (j <= i) continue;
                const double c = trig_cos_sep(&rc->trig[i], &rc->trig[j]);
                st.num_sightline_pairs_considered++;
                if (c <= rc->cos_max) continue;
                if (nc == cap) { cap *= 2; cs = (cand *)realloc(cs, sizeof(cand) * cap); }
                cs[nc].j = j; cs[nc].key = rc->sky->bin_of_los[j]; cs[nc].c = c; ++nc;
            }
            qsort(cs, nc, sizeof(cand), cmp_cand);
        } else {
            /* PlateBins::getBinIndicesWithinRadius (PlateBins.h:74-85): plates
             * in ascending id whose centre lies in the annulus
             * |acos(c) - 1.5 deg| < radius; members j > i of existing bins. */
            for (int64_t p = 0; p < rc->sky->n_plates; ++p) {
                const double cs_ = trig_cos_sep(&rc->trig[i], &rc->plate_trig[p]);
                const double sep = acos(cs_);
                if (!(fabs(sep - 1.5 * deg2rad) < rc->max_ang)) continue;
                const int64_t kk = find_key(bi, rc->sky->plate_id[p]);
                if (kk < 0) continue;
                for (int64_t mm = bi->member_ptr[kk]; mm < bi->member_ptr[kk + 1]; ++mm) {
                    const int64_t j = bi->members[mm];
                    if (j <= i) continue;
                    const double c = trig_cos_sep(&rc->trig[i], &rc->trig[j]);
                    st.num_sightline_pairs_considered++;
                    if (c <= rc->cos_max) continue;
                    if (nc == cap) { cap *= 2; cs = (cand *)realloc(cs, sizeof(cand) * cap); }
                    cs[nc].j = j; cs[nc].key = rc->sky->plate_id[p]; cs[nc].c = c; ++nc;
                }
            }
        }
        for (size_t q = 0; q < nc; ++q) {
            const int64_t j = cs[q].j;
            const int64_t nj = s->offsets[j + 1] - s->offsets[j];
            st.num_sightline_pairs_used++;
            st.num_pixel_pairs += (uint64_t)(ni * nj);
            st.num_pixel_pairs_used += ora_accumulate_los_pair(s, rc->g, i, j, cs[q].c, xi);
        }
    }
    free(cs);
    rc->task_stats[t] = st;
}

static void *worker(void *arg)
{
    run_ctx *rc = (run_ctx *)arg;
    for (;;) {
        int64_t t = __atomic_fetch_add(&rc->next, 1, __ATOMIC_RELAXED);
        if (t >= rc->n_tasks) break;
        run_task(rc, t);
    }
    return NULL;
}

static const double *g_dec_for_sort;
static int cmp_by_dec(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    double dx = g_dec_for_sort[x], dy = g_dec_for_sort[y];
    if (dx != dy) return (dx > dy) - (dx < dy);
    return (x > y) - (x < y);
}

int64_t ora_xi_run(const ora_sightlines *s, const ora_grid *g, double scale,
                   const ora_skybins *sky, int nthreads,
                   int64_t task_stride, int64_t task_phase,
                   ora_xibin *xi_final, ora_xibin *xi_raw,
                   ora_xibin *per_skybin, int32_t *skybin_keys,
                   ora_stats *stats)
{
    if (!s || !g || !sky || s->n_los < 0) return -1;
    if (task_stride < 1) task_stride = 1;
    const uint32_t nbins = (uint32_t)g->nbins[0] * (uint32_t)g->nbins[1] * (uint32_t)g->nbins[2];
    if (nbins == 0) return -1;
    if (nthreads < 1) nthreads = 1;

    bin_index bi;
    if (build_bin_index(sky->bin_of_los, s->n_los, &bi) != 0) return -1;

    sky_trig *trig = (sky_trig *)malloc(sizeof(sky_trig) * (size_t)(s->n_los + 1));
    for (int64_t i = 0; i < s->n_los; ++i) trig[i] = make_trig(s->ra[i], s->dec[i]);
    sky_trig *ptrig = NULL;
    if (sky->kind == ORA_SKY_PLATES) {
        ptrig = (sky_trig *)malloc(sizeof(sky_trig) * (size_t)(sky->n_plates + 1));
        /* the plate list is a std::map: iterate in ascending plate id */
        int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)(sky->n_plates + 1));
        for (int64_t p = 0; p < sky->n_plates; ++p) ord[p] = p;
        for (int64_t a = 1; a < sky->n_plates; ++a) {   /* insertion sort, small lists */
            int64_t v = ord[a], b = a - 1;
            while (b >= 0 && sky->plate_id[ord[b]] > sky->plate_id[v]) { ord[b + 1] = ord[b]; --b; }
            ord[b + 1] = v;
        }
        /* sorted copies */
        int32_t *pid = (int32_t *)malloc(sizeof(int32_t) * (size_t)(sky->n_plates + 1));
        for (int64_t p = 0; p < sky->n_plates; ++p) {
            pid[p] = sky->plate_id[ord[p]];
            ptrig[p] = make_trig(sky->plate_ra[ord[p]], sky->plate_dec[ord[p]]);
        }
        free(ord);
        /* keep the sorted ids alive for the run via a shallow copy of sky */
        ora_skybins *sky2 = (ora_skybins *)malloc(sizeof(ora_skybins));
        *sky2 = *sky;
        sky2->plate_id = pid;
        sky = sky2;
    }

    int64_t *by_dec = (int64_t *)malloc(sizeof(int64_t) * (size_t)(s->n_los + 1));
    double *dec_sorted = (double *)malloc(sizeof(double) * (size_t)(s->n_los + 1));
    for (int64_t i = 0; i < s->n_los; ++i) by_dec[i] = i;
    g_dec_for_sort = s->dec;
    qsort(by_dec, (size_t)s->n_los, sizeof(int64_t), cmp_by_dec);
    for (int64_t i = 0; i < s->n_los; ++i) dec_sorted[i] = s->dec[by_dec[i]];

    /* tasks: occupied bins in first-insertion order (XiEstimator.cc:37-40) */
    int64_t n_tasks = 0;
    int64_t *task_key_idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(bi.n_keys + 1));
    for (int64_t t = 0; t < bi.n_keys; ++t)
        if (t % task_stride == task_phase)
            task_key_idx[n_tasks++] = find_key(&bi, bi.keys_insertion[t]);

    run_ctx rc;
    memset(&rc, 0, sizeof(rc));
    rc.s = s; rc.g = g; rc.sky = sky; rc.bi = &bi; rc.trig = trig; rc.plate_trig = ptrig;
    rc.by_dec = by_dec; rc.dec_sorted = dec_sorted;
    rc.max_ang = ora_max_angular_scale(g, scale);      /* XiEstimator.cc:24 */
    rc.cos_max = cos(ora_max_angular_scale(g, scale)); /* XiEstimator.cc:25 */
    rc.nbins = nbins;
    rc.n_tasks = n_tasks;
    rc.task_key_idx = task_key_idx;
    rc.task_xi = (ora_xibin *)malloc(sizeof(ora_xibin) * (size_t)nbins * (size_t)(n_tasks > 0 ? n_tasks : 1));
    rc.task_stats = (ora_stats *)calloc((size_t)(n_tasks > 0 ? n_tasks : 1), sizeof(ora_stats));
    rc.next = 0;

    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &rc);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);

    /* merge in ascending key order (std::map iteration, XiEstimator.cc:298-305) */
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_tasks + 1));
    {
        int64_t o = 0;
        for (int64_t k = 0; k < bi.n_keys; ++k)
            for (int64_t t = 0; t < n_tasks; ++t)
                if (task_key_idx[t] == k) { order[o++] = t; break; }
    }
    ora_stats tot;
    memset(&tot, 0, sizeof(tot));
    for (int64_t t = 0; t < n_tasks; ++t) {
        tot.num_sightline_pairs_used += rc.task_stats[t].num_sightline_pairs_used;
        tot.num_pixel_pairs += rc.task_stats[t].num_pixel_pairs;
        tot.num_pixel_pairs_used += rc.task_stats[t].num_pixel_pairs_used;
        tot.num_pixels += rc.task_stats[t].num_pixels;
        tot.num_sightline_pairs_considered += rc.task_stats[t].num_sightline_pairs_considered;
    }
    if (stats) *stats = tot;
    for (uint32_t b = 0; b < nbins; ++b) {
        ora_xibin acc;
        memset(&acc, 0, sizeof(acc));
        for (int64_t o = 0; o < n_tasks; ++o)
            xibin_add(&acc, &rc.task_xi[(size_t)order[o] * nbins + b]);
        if (xi_raw) xi_raw[b] = acc;
        if (xi_final) { ora_finalize(&acc); xi_final[b] = acc; }
    }
    if (per_skybin || skybin_keys) {
        for (int64_t o = 0; o < n_tasks; ++o) {
            const int64_t t = order[o];
            if (skybin_keys) skybin_keys[o] = bi.keys_sorted[task_key_idx[t]];
            if (per_skybin)
                memcpy(per_skybin + (size_t)o * nbins, rc.task_xi + (size_t)t * nbins,
                       sizeof(ora_xibin) * nbins);
        }
    }

    free(order);
    free(rc.task_xi);
    free(rc.task_stats);
    free(task_key_idx);
    free(by_dec);
    free(dec_sorted);
    free(trig);
    if (ptrig) {
        free(ptrig);
        free((void *)sky->plate_id);
        free((void *)sky);
    }
    free_bin_index(&bi);
    return n_tasks;
}
