// tos_brute.hip -- 3D isotropic brute-force pair counts: the FP64
// BinPair semantics (old/BinPair.h:28-47) and the FP32 semantics of the
// reference GPU kernel (gpu/brute.cu:26-55), all N(N-1)/2 pairs.
//
// Lanes hold 64 points i of a block; the loop walks points j > i of a run,
// wave-uniform, four at a time through the scalar cache (one s_load of four
// packed points).  A workgroup keeps `copies` interleaved copies of its LDS
// histogram ([slot][copy]) and lane l adds into copy l mod copies: with 50
// bins, 64 lanes adding into one copy collide on the same addresses most of
// the time, and those collisions serialise in the LDS; with 16 copies the
// lanes sharing a copy (l, l+16, l+32, l+48) fall in different passes of a
// 64-bit LDS access.  The copies are summed in a fixed order at the end,
// then written as the workgroup's partial row.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tos_internal.h"

namespace tos {
namespace {

constexpr int kWave = 64;
// 8 wavefronts share a workgroup's histogram copies (measured: 512 threads
// 11% faster than 256 at 10^6 points; 64 copies, one per lane, slower)
constexpr int kBruteBlock = 512;

template <int MODE, bool POW2>
__global__ void __launch_bounds__(kBruteBlock) brute3d_kernel(BruteArgs a)
{
    extern __shared__ double lds_raw[];
    const int nslot = (MODE == TOS_BRUTE_FP64) ? a.nbins : a.nbins + 2;
    const int H = a.copies;
    double *hd = lds_raw;                                                       // [nslot][H]
    double *hw = lds_raw + (size_t)nslot * H;                                   // [nslot][H]
    uint32_t *hc = reinterpret_cast<uint32_t *>(lds_raw + (size_t)2 * nslot * H);   // [nslot][H]
    for (int b = threadIdx.x; b < nslot * H; b += blockDim.x) {
        hd[b] = 0.0;
        hw[b] = 0.0;
        hc[b] = 0u;
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nwave = blockDim.x / kWave;
    const int copy = lane & (H - 1);
    // BinPair ctor (old/BinPair.h:18-23) / brute.cu:45
    const double bin_width = (a.mx - a.mn) / (double)a.nbins;
    const double min_sq = a.mn * a.mn, max_sq = a.mx * a.mx;
    const float mnf = (float)a.mn, mxf = (float)a.mx;
    const float spacing = (mxf - mnf) / (float)a.nbins;

    for (int32_t t = blockIdx.x * nwave + wave; t < a.n_tiles; t += gridDim.x * nwave) {
        const int32_t ib = a.tiles[2 * t], jb = a.tiles[2 * t + 1];
        const int32_t je = (int32_t)min<int64_t>(a.n, (int64_t)jb + a.chunk);
        const int32_t i = ib + lane;
        const bool active = i < a.n;
        const int32_t ii = active ? i : 0;
        const double4 pi = a.pts[ii];
        const double wi = a.w[ii];
        const float xf = (float)pi.x, yf = (float)pi.y, zf = (float)pi.z, df = (float)pi.w, wf = (float)wi;

        auto pair = [&](const double4 &pj, const double wj, const int32_t j) {
            const bool pair_ok = active & (j > i);
            int b = -1;
            double term = 0.0, wgt = 0.0;
            if constexpr (MODE == TOS_BRUTE_FP64) {
                const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const double rsq = dx * dx + dy * dy + dz * dz;
                if (!(rsq < min_sq || rsq >= max_sq)) {
                    // BinPair: floor((sqrt(rsq) - min) / binWidth); a power-of-two
                    // width divides exactly as a multiplication by its inverse
                    const double u = POW2 ? (sqrt(rsq) - a.mn) * a.inv_width : (sqrt(rsq) - a.mn) / bin_width;
                    const double idx_d = floor(u);
                    if (idx_d >= 0.0 && idx_d < (double)a.nbins) b = (int)idx_d;
                }
                wgt = wi * wj;
                term = wgt * (pi.w * pj.w);
            } else {
                const float dx = xf - (float)pj.x, dy = yf - (float)pj.y, dz = zf - (float)pj.z;
                const float s2 = dx * dx + dy * dy + dz * dz;
                const float sep = sqrtf(s2);
                const float fw = wf * (float)wj;
                if (sep < mnf) b = 0;
                else if (sep >= mxf) b = a.nbins + 1;
                else b = (int)floorf(POW2 ? (sep - mnf) * a.inv_spacing : (sep - mnf) / spacing) + 1;
                if (b > a.nbins + 1) b = a.nbins + 1;
                const float wd = fw * df;
                term = (double)(wd * (float)pj.w);
                wgt = (double)fw;
            }
            if (pair_ok && b >= 0) {
                const int s = b * H + copy;
                __hip_atomic_fetch_add(&hc[s], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                unsafeAtomicAdd(&hd[s], term);
                unsafeAtomicAdd(&hw[s], wgt);
            }
        };

        int32_t j = jb;
        for (; j + 4 <= je; j += 4) {
            const double4 p0 = a.pts[j], p1 = a.pts[j + 1], p2 = a.pts[j + 2], p3 = a.pts[j + 3];
            const double w0 = a.w[j], w1 = a.w[j + 1], w2 = a.w[j + 2], w3 = a.w[j + 3];
            pair(p0, w0, j);
            pair(p1, w1, j + 1);
            pair(p2, w2, j + 2);
            pair(p3, w3, j + 3);
        }
        for (; j < je; ++j) pair(a.pts[j], a.w[j], j);
    }
    __syncthreads();
    // copies summed in copy order, then one partial row per workgroup
    const size_t base = (size_t)blockIdx.x * nslot;
    for (int s = threadIdx.x; s < nslot; s += blockDim.x) {
        uint32_t c = 0;
        double sd = 0.0, sw = 0.0;
        for (int k = 0; k < H; ++k) {
            c += hc[s * H + k];
            sd += hd[s * H + k];
            sw += hw[s * H + k];
        }
        a.part_cnt[base + s] = c;
        a.part_sum[(base + s) * 2] = sd;
        a.part_sum[(base + s) * 2 + 1] = sw;
    }
}

template <int MODE, bool POW2>
hipError_t launch_mode(const BruteArgs &a, int n_wg, size_t lds, hipStream_t st)
{
    auto fn = brute3d_kernel<MODE, POW2>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(n_wg), dim3(kBruteBlock), lds, st, a);
    return hipGetLastError();
}

}  // namespace

size_t brute3d_lds_bytes(int mode, int32_t nbins, int copies)
{
    const size_t nslot = mode == TOS_BRUTE_FP64 ? (size_t)nbins : (size_t)nbins + 2;
    return nslot * (size_t)copies * 20;
}

int launch_brute3d(int mode, bool pow2, const BruteArgs &a, int n_wg, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const size_t lds = brute3d_lds_bytes(mode, a.nbins, a.copies);
    hipError_t e;
    if (mode == TOS_BRUTE_FP64)
        e = pow2 ? launch_mode<TOS_BRUTE_FP64, true>(a, n_wg, lds, st) : launch_mode<TOS_BRUTE_FP64, false>(a, n_wg, lds, st);
    else
        e = pow2 ? launch_mode<TOS_BRUTE_FP32, true>(a, n_wg, lds, st) : launch_mode<TOS_BRUTE_FP32, false>(a, n_wg, lds, st);
    return (int)e;
}

}  // namespace tos
