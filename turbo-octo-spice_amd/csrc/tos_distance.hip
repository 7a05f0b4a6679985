// tos_distance.hip -- the per-pixel distance transform of h5healxi.cc:196-216
// on the device: one thread per pixel evaluates the FP64 line-of-sight
// comoving distance of its wavelength (tos_cosmo.h, the same arithmetic as
// the host's tos_cosmo_distances) and rounds it to the pixel's float.
#include <hip/hip_runtime.h>

#include "tos_capi_impl.h"
#include "tos_cosmo.h"

using namespace tos;

namespace tos {
namespace {

__global__ void __launch_bounds__(256) pixel_distance_kernel(const float *__restrict__ loglam, int64_t n, double ol,
                                                             double om, double log_lya, float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = pixel_distance(ol, om, log_lya, loglam[i]);
}

}  // namespace
}  // namespace tos

extern "C" int tos_pixel_distances(tos_ctx *ctx, double omega_lambda, double omega_matter, int64_t n,
                                   const float *loglam, float *distance)
{
    if (!ctx) return TOS_ERR_INVALID;
    if (n < 0 || (n > 0 && (!loglam || !distance)) || !std::isfinite(omega_lambda) || !std::isfinite(omega_matter)) {
        set_err(ctx, "tos_pixel_distances: invalid argument");
        return TOS_ERR_INVALID;
    }
    if (n == 0) return TOS_OK;
    if (omega_matter == 0) omega_matter = 1 - omega_lambda;   // h5healxi.cc:124
    hipSetDevice(ctx->device);
    const size_t nb = (size_t)n * sizeof(float);
    int rc;
    if ((rc = upload(ctx, ctx->scratch_a, loglam, nb, "pixel loglam")) != TOS_OK) return rc;
    if (ctx->scratch_b.reserve(nb) != TOS_OK) {
        set_err(ctx, "tos_pixel_distances: device allocation failed");
        return TOS_ERR_NOMEM;
    }
    const double log_lya = std::log10(1216.0);   // constants.h:9-10 (lyA, logLyA)
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffff) {
        set_err(ctx, "tos_pixel_distances: too many pixels for one launch");
        return TOS_ERR_INVALID;
    }
    hipLaunchKernelGGL(pixel_distance_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       (const float *)ctx->scratch_a.ptr, n, omega_lambda, omega_matter, log_lya,
                       (float *)ctx->scratch_b.ptr);
    hipError_t he = hipGetLastError();
    if (he == hipSuccess) he = hipMemcpyAsync(distance, ctx->scratch_b.ptr, nb, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "tos_pixel_distances");
    return TOS_OK;
}
