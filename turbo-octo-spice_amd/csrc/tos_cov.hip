// tos_cov.hip -- weighted covariance of the sky-bin subsamples
// (XiEstimator.cc:60-97: every sky-bin histogram is one sample
// xi_b = didj_b / wgt_b over the bins with wgt_b > 0, weighted by the sum of
// those wgt_b; likely::CovarianceAccumulator then gives the weighted
// covariance of the samples).
//
//   w_k  = sum_b [wgt_kb > 0] wgt_kb            x_kb = wgt_kb > 0 ? didj_kb / wgt_kb : 0
//   mu_b = sum_k w_k x_kb / W,  W = sum_k w_k
//   C_ab = sum_k w_k (x_ka - mu_a)(x_kb - mu_b) / W
//
// C is a K x N by K x N contraction (K samples, N bins: ~2000 x 2500 at
// full-survey scale), so it runs on the FP64 matrix cores: 16x16x4 MFMA
// tiles, four wavefronts per 32x32 block of C.  The centring and weighting
// are done once, in the pass that forms the two K x N operands.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tos_internal.h"

namespace tos {
namespace {

typedef double double4_v __attribute__((ext_vector_type(4)));

// per sample: x_k (in place of the operand rows) and w_k; one workgroup per
// sample, bins summed in a fixed tree order
__global__ void __launch_bounds__(256) cov_samples(const double *didj, const double *wgt, int64_t k_n, int32_t n,
                                                   double *x, double *w)
{
    const int64_t k = blockIdx.x;
    if (k >= k_n) return;
    __shared__ double part[256];
    double s = 0.0;
    for (int32_t b = threadIdx.x; b < n; b += blockDim.x) {
        const double g = wgt[k * n + b];
        const bool on = g > 0;
        x[k * n + b] = on ? didj[k * n + b] / g : 0.0;
        if (on) s += g;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) w[k] = part[0];
}

// mean over samples per bin, in sample order; wsum[0] = W (same order)
__global__ void __launch_bounds__(256) cov_mean(const double *x, const double *w, int64_t k_n, int32_t n, double *mu,
                                                double *wsum)
{
    const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    double W = 0.0;
    for (int64_t k = 0; k < k_n; ++k) W += w[k];
    if (b == 0) wsum[0] = W;
    if (b >= n) return;
    double s = 0.0;
    for (int64_t k = 0; k < k_n; ++k) s += w[k] * x[k * n + b];
    mu[b] = W > 0 ? s / W : 0.0;
}

// operands: A_kb = w_k (x_kb - mu_b), B_kb = x_kb - mu_b (B in place of x)
__global__ void __launch_bounds__(256) cov_center(double *x, const double *w, const double *mu, int64_t k_n, int32_t n,
                                                  double *a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k_n * n) return;
    const int64_t k = i / n;
    const int32_t b = (int32_t)(i - k * n);
    const double d = x[i] - mu[b];
    x[i] = d;
    a[i] = w[k] * d;
}

// C = A^T B / W on the matrix cores.  Block (bx, by): C rows [32 bx, +32),
// columns [32 by, +32); wavefront q owns the 16x16 quadrant (q / 2, q % 2).
// v_mfma_f64_16x16x4f64: lane l supplies A[k + l/16][i0 + l%16] and
// B[k + l/16][j0 + l%16]; it holds D[l/16 + 4 r][l%16], r = 0..3 (lane
// layout checked on the device: tests/test_gpu_covariance.py).
__global__ void __launch_bounds__(256) cov_gemm(const double *A, const double *B, const double *wsum, int64_t k_n,
                                                int32_t n, double *C)
{
    const int lane = threadIdx.x & 63;
    const int q = threadIdx.x >> 6;
    const int32_t i0 = blockIdx.x * 32 + 16 * (q >> 1);
    const int32_t j0 = blockIdx.y * 32 + 16 * (q & 1);
    const int32_t ci = i0 + (lane & 15), cj = j0 + (lane & 15);
    const int kk = lane >> 4;
    const bool iok = ci < n, jok = cj < n;
    double4_v acc = {0.0, 0.0, 0.0, 0.0};
    int64_t k = 0;
    for (; k + 16 <= k_n; k += 16) {
        double av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t row = k + 4 * u + kk;
            av[u] = iok ? A[row * n + ci] : 0.0;
            bv[u] = jok ? B[row * n + cj] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
    }
    for (; k < k_n; k += 4) {
        const int64_t row = k + kk;
        const bool kok = row < k_n;
        const double av = (iok && kok) ? A[row * n + ci] : 0.0;
        const double bv = (jok && kok) ? B[row * n + cj] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
    const double W = wsum[0];
    const double inv = W > 0 ? 1.0 / W : 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int32_t oi = i0 + kk + 4 * r;
        if (oi < n && jok) C[(int64_t)oi * n + cj] = acc[r] * inv;
    }
}

}  // namespace

int launch_covariance(const double *didj, const double *wgt, int64_t k_n, int32_t n, double *x, double *a,
                      double *w, double *mu, double *wsum, double *cov, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (k_n <= 0 || n <= 0) return 0;
    hipLaunchKernelGGL(cov_samples, dim3((unsigned)k_n), dim3(256), 0, st, didj, wgt, k_n, n, x, w);
    hipLaunchKernelGGL(cov_mean, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, w, k_n, n, mu, wsum);
    const int64_t tot = k_n * n;
    hipLaunchKernelGGL(cov_center, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, w, mu, k_n, n, a);
    const unsigned tiles = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL(cov_gemm, dim3(tiles, tiles), dim3(256), 0, st, a, x, wsum, k_n, n, cov);
    return (int)hipGetLastError();
}

}  // namespace tos
