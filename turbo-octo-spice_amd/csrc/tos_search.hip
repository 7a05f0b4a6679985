// tos_search.hip -- candidate sightline pairs on the GPU (the skybin_xi_task
// search, XiEstimator.cc:121-159, for HEALPix sky bins).
//
// One wavefront per primary sightline i, primaries in the order of the
// candidate rows (occupied bins, then members).  The lanes walk the members
// of i's neighbour bins 64 at a time; a member j > i is a candidate, kept
// when cos(sep) > cos(max_ang) with cos(sep) in the reference's operation
// order (types.h:37-39; this file is built with -ffp-contract=off, so the
// FP64 products and sums round exactly as on the host).  A first pass counts
// each row; the second writes the kept partners in member order with a
// ballot-ordered compaction, so the rows equal the host search's.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tos_internal.h"

namespace tos {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ double cos_sep_dev(const double4 a, const double4 b)
{
    const double t1 = a.x * b.x;
    const double t2 = a.y * b.y;
    const double t3 = a.z * b.z;
    const double t4 = a.w * b.w;
    return t1 + t2 * (t3 + t4);
}

template <bool FILL>
__global__ void __launch_bounds__(256) search_kernel(SearchArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (r >= a.n_prim) return;   // whole wavefronts leave together
    const int32_t i = a.prim[r];
    const int32_t cell = a.prim_cell[r];
    const double4 ti = a.trig[i];
    const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    int64_t kept = 0;                 // wave-uniform
    uint64_t considered = 0, partner_pixels = 0;
    int64_t out = FILL ? a.offset[r] : 0;
    // the fill pass repeats the count pass's arithmetic; the row's end
    // bounds its writes all the same
    const int64_t end = FILL ? out + a.count[r] : 0;
    for (int32_t q = a.nbr_ptr[cell]; q < a.nbr_ptr[cell + 1]; ++q) {
        const int32_t b = a.nbr[q];
        const int32_t m1 = a.cell_ptr[b + 1];
        for (int32_t k0 = a.cell_ptr[b]; k0 < m1; k0 += kWave) {
            const int32_t k = k0 + lane;
            const bool valid = k < m1;
            const int32_t j = valid ? a.members[k] : -1;
            const bool cand = valid && j > i;
            double c = 0.0;
            if (cand) c = cos_sep_dev(ti, a.trig[j]);
            const bool keep = cand && (c > a.cos_max);
            const uint64_t mk = __ballot(keep);
            const uint64_t mc = __ballot(cand);
            considered += (uint64_t)__popcll(mc);
            if constexpr (FILL) {
                const int64_t pos = out + kept + __popcll(mk & lanes_below);
                if (keep && pos < end) {
                    a.other[pos] = j;
                    a.csep[pos] = c;
                }
            } else {
                uint32_t nj = keep ? (uint32_t)a.los[j].n : 0u;
                // wave sum of the kept partners' pixel counts
                for (int off = 32; off > 0; off >>= 1) nj += __shfl_xor(nj, off);
                partner_pixels += nj;
            }
            kept += __popcll(mk);
        }
    }
    if constexpr (!FILL) {
        if (lane == 0) {
            a.count[r] = (int32_t)kept;
            a.considered[r] = considered;
            a.pixel_pairs[r] = partner_pixels * (uint64_t)a.los[i].n;
        }
    }
}

}  // namespace

int launch_search(const SearchArgs &a, bool fill, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int waves_per_block = 4;
    const int64_t blocks = (a.n_prim + waves_per_block - 1) / waves_per_block;
    if (blocks <= 0) return 0;
    if (fill)
        hipLaunchKernelGGL(search_kernel<true>, dim3((unsigned)blocks), dim3(64 * waves_per_block), 0, st, a);
    else
        hipLaunchKernelGGL(search_kernel<false>, dim3((unsigned)blocks), dim3(64 * waves_per_block), 0, st, a);
    return (int)hipGetLastError();
}

}  // namespace tos
