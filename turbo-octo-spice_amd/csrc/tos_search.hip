// tos_search.hip -- candidate sightline pairs on the GPU (the skybin_xi_task
// search, XiEstimator.cc:121-159, for HEALPix and plate sky bins).
//
// One wavefront per primary sightline i, primaries in the order of the
// candidate rows (occupied bins, then members).  The lanes walk the members
// of i's neighbour bins 64 at a time (HEALPix: the host's neighbour-cell
// lists; plates: the annulus test per plate, 64 plates at a time); a member
// j > i is a candidate, kept
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

// Members j > i of occupied bin b, 64 at a time: the count pass sums kept
// partners, candidates and partner pixels; the fill pass writes the kept
// partners in member order.
template <bool FILL>
__device__ __forceinline__ void visit_bin(const SearchArgs &a, int32_t b, int32_t i, const double4 ti, int lane,
                                          uint64_t lanes_below, int64_t out, int64_t end, int64_t &kept,
                                          uint64_t &considered, uint64_t &partner_pixels)
{
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

template <bool FILL, bool PLATES>
__global__ void __launch_bounds__(256) search_kernel(SearchArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (r >= a.n_prim) return;   // whole wavefronts leave together
    const int32_t i = a.prim[r];
    const double4 ti = a.trig[i];
    const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    int64_t kept = 0;                 // wave-uniform
    uint64_t considered = 0, partner_pixels = 0;
    int64_t out = FILL ? a.offset[r] : 0;
    // the fill pass repeats the count pass's arithmetic; the row's end
    // bounds its writes all the same
    const int64_t end = FILL ? out + a.count[r] : 0;
    if constexpr (PLATES) {
        // PlateBins::getBinIndicesWithinRadius (PlateBins.h:74-85): plates in
        // ascending id whose centre is in the annulus, tested on c against
        // thresholds; the plates of one 64-plate chunk are visited in order
        for (int32_t p0 = 0; p0 < a.n_plate; p0 += kWave) {
            const int32_t p = p0 + lane;
            bool in = false;
            if (p < a.n_plate && a.plate_bin[p] >= 0) {
                const double c = cos_sep_dev(ti, a.plate_trig[p]);
                // c outside [-1, 1] is a NaN angle on the host: never in
                in = c >= -1.0 && c <= 1.0;
                if (a.has_up) in = in && c > a.c_up + a.margin;
                if (a.has_lo) in = in && c < a.c_lo - a.margin;
                // the bounds themselves carry the host's rounding of cos:
                // a c this close to one may fall on either side of acos's
                const bool border = (a.has_up && fabs(c - a.c_up) <= a.margin) ||
                                    (a.has_lo && fabs(c - a.c_lo) <= a.margin);
                if (!FILL && border) atomicOr(a.undecided, 1);
            }
            uint64_t mp = __ballot(in);
            while (mp) {
                const int q = __ffsll((long long)mp) - 1;
                mp &= mp - 1;
                const int32_t b = a.plate_bin[p0 + q];
                visit_bin<FILL>(a, b, i, ti, lane, lanes_below, out, end, kept, considered, partner_pixels);
            }
        }
    } else {
        const int32_t cell = a.prim_cell[r];
        for (int32_t q = a.nbr_ptr[cell]; q < a.nbr_ptr[cell + 1]; ++q)
            visit_bin<FILL>(a, a.nbr[q], i, ti, lane, lanes_below, out, end, kept, considered, partner_pixels);
    }
    if constexpr (!FILL) {
        if (lane == 0) {
            a.count[r] = (int32_t)kept;
            a.considered[r] = considered;
            a.pixel_pairs[r] = partner_pixels * (uint64_t)a.los[i].n;
        }
    }
}

// Work items of row r, the host's build_items rules: partner runs of at
// most step_cap partner pixels for every 64-pixel lane chunk of the
// primary, then, for a short remainder, transposed runs costed by partner
// chunks.  The count pass (FILL false) only counts.
template <bool FILL>
__global__ void __launch_bounds__(256) items_kernel(ItemArgs a)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n_rows) return;
    const int32_t i = a.row_los[r];
    const int32_t ni = a.los[i].n;
    const int64_t p0 = a.row_ptr[r], p1 = a.row_ptr[r + 1];
    const int32_t full = ni / 64, rem = ni % 64;
    const bool transpose_tail = rem > 0 && rem < a.transpose_below;
    const int32_t lane_chunks = full + ((rem > 0 && !transpose_tail) ? 1 : 0);
    int64_t out = FILL ? a.row_item_ptr[r] : 0;
    const int64_t end = FILL ? a.row_item_ptr[r + 1] : 0;
    int32_t count = 0;
    auto emit = [&](XiItem it, uint64_t cost) {
        if constexpr (FILL) {
            if (out < end) {
                a.items[out] = it;
                a.item_cost[out] = cost;
            }
        }
        ++out;
        ++count;
    };
    int64_t s0 = p0;
    while (s0 < p1 && lane_chunks > 0) {
        int64_t b = s0;
        int64_t steps = 0;
        while (b < p1 && (b == s0 || steps + a.los[a.other[b]].n <= a.step_cap)) {
            steps += a.los[a.other[b]].n;
            ++b;
        }
        for (int32_t c = 0; c < lane_chunks; ++c) emit(XiItem{i, 64 * c, (int32_t)s0, (int32_t)b}, (uint64_t)steps + 8);
        s0 = b;
    }
    if (transpose_tail) {
        s0 = p0;
        while (s0 < p1) {
            int64_t b = s0;
            int64_t steps = 0;
            while (b < p1) {
                const int64_t cst = (int64_t)((a.los[a.other[b]].n + 63) / 64) * (rem + 2);
                if (b != s0 && steps + cst > a.step_cap) break;
                steps += cst;
                ++b;
            }
            emit(XiItem{i, (64 * full) | kTransposed, (int32_t)s0, (int32_t)b}, (uint64_t)steps + 8);
            s0 = b;
        }
    }
    if constexpr (!FILL) a.count[r] = count;
}

}  // namespace

int launch_items(const ItemArgs &a, bool fill, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int64_t blocks = (a.n_rows + 255) / 256;
    if (blocks <= 0) return 0;
    if (fill) hipLaunchKernelGGL(items_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(items_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

int launch_search(const SearchArgs &a, bool fill, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int waves_per_block = 4;
    const int64_t blocks = (a.n_prim + waves_per_block - 1) / waves_per_block;
    if (blocks <= 0) return 0;
    const dim3 grid((unsigned)blocks), block(64 * waves_per_block);
    if (a.plates) {
        if (fill) hipLaunchKernelGGL((search_kernel<true, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((search_kernel<false, true>), grid, block, 0, st, a);
    } else {
        if (fill) hipLaunchKernelGGL((search_kernel<true, false>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((search_kernel<false, false>), grid, block, 0, st, a);
    }
    return (int)hipGetLastError();
}

}  // namespace tos
