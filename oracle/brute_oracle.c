/*
 * brute_oracle.c -- 3D isotropic pair counting (bruteGPU / BinPair).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * ora_brute3d_f64 restates BinPair::binPair (turbooctospice/old/BinPair.h:28-47)
 * over every pair i < j in the order of BrutePairSearch
 * (turbooctospice/old/BrutePairSearch.h:31-50), with the pair arithmetic of
 * Pair::separationSq/weight/product (turbooctospice/old/Pair.h:26-41).
 * ora_brute3d_f32 restates the FP32 separation and bin rule of histo_kernel
 * (gpu/brute.cu:26-55); the reference sums with float atomics in an
 * unspecified order, here the per-pair float terms are summed in double.
 *
 * Pinning: sample-data/delta-nanoer.dat (a 2x2x2 lattice at 5 Mpc/h spacing:
 * 12 pairs at 5, 12 at 5*sqrt2, 4 at 5*sqrt3) counted by hand; a numpy
 * float64 all-pairs histogram on delta-nano.dat away from bin edges;
 * sum(counts) == number of pairs with min <= r < max.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

void ora_brute3d_f64(const double *x, const double *y, const double *z,
                     const double *d, const double *w, int64_t n,
                     double min, double max, int32_t nbins,
                     double *dsum, double *wsum, int64_t *counts)
{
    memset(dsum, 0, sizeof(double) * (size_t)nbins);
    memset(wsum, 0, sizeof(double) * (size_t)nbins);
    memset(counts, 0, sizeof(int64_t) * (size_t)nbins);
    /* BinPair ctor (BinPair.h:18-23); _nBins is stored as a double */
    const double bin_width = (max - min) / (double)nbins;
    const double min_sq = min * min;
    const double max_sq = max * max;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = i + 1; j < n; ++j) {
            const double dx = x[i] - x[j];
            const double dy = y[i] - y[j];
            const double dz = z[i] - z[j];
            const double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < min_sq || rsq >= max_sq) continue;
            const double idx_d = floor((sqrt(rsq) - min) / bin_width);
            /* sqrt can round up to `max` (or below `min`): the reference
             * writes out of bounds there; such pairs are skipped */
            if (!(idx_d >= 0.0 && idx_d < (double)nbins)) continue;
            const int idx = (int)idx_d;
            const double wgt = w[i] * w[j];
            wsum[idx] += wgt;
            dsum[idx] += wgt * (d[i] * d[j]);
            counts[idx] += 1;
        }
    }
}

void ora_brute3d_f32(const float *x, const float *y, const float *z,
                     const float *d, const float *w, int64_t n,
                     float min, float max, int32_t nbins,
                     double *dsum, double *wsum, int64_t *counts)
{
    const int32_t nh = nbins + 2;
    memset(dsum, 0, sizeof(double) * (size_t)nh);
    memset(wsum, 0, sizeof(double) * (size_t)nh);
    memset(counts, 0, sizeof(int64_t) * (size_t)nh);
    const float spacing = (max - min) / (float)nbins;   /* brute.cu:45 */
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = i + 1; j < n; ++j) {
            const float dx = x[i] - x[j];
            const float dy = y[i] - y[j];
            const float dz = z[i] - z[j];
            const float s2 = dx * dx + dy * dy + dz * dz;
            const float sep = sqrtf(s2);
            const float wgt = w[i] * w[j];
            int idx;
            if (sep < min) idx = 0;
            else if (sep >= max) idx = nbins + 1;
            else idx = (int)floorf((sep - min) / spacing) + 1;
            if (idx > nbins + 1) idx = nbins + 1;
            const float wd = wgt * d[i];
            const float term = wd * d[j];
            dsum[idx] += (double)term;
            wsum[idx] += (double)wgt;
            counts[idx] += 1;
        }
    }
}
