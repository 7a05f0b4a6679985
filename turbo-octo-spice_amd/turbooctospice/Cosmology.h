// Cosmology.h -- the LCDM distances used by the h5healxi distance transform
// (h5healxi.cc:123-131, 196-239; cosmo::LambdaCdmUniverse is not vendored
// by the reference).  Distances in Mpc/h.
#ifndef TOS_AMD_TURBOOCTOSPICE_COSMOLOGY
#define TOS_AMD_TURBOOCTOSPICE_COSMOLOGY

#include <cstdint>
#include <string>

#include "RuntimeError.h"
#include "tos_amd.h"

namespace turbooctospice {

class LambdaCdmUniverse {
public:
    /// OmegaMatter = 0 means 1 - OmegaLambda (h5healxi.cc:124).
    LambdaCdmUniverse(double omega_lambda, double omega_matter)
        : ol_(omega_lambda), om_(omega_matter == 0 ? 1 - omega_lambda : omega_matter)
    {
        if (!(om_ > 0)) throw RuntimeError("LambdaCdmUniverse: Omega_matter must be positive");
    }
    double getLineOfSightComovingDistance(double z) const
    {
        double dc = 0;
        if (tos_cosmo_distances(ol_, om_, 1, &z, &dc, nullptr) != TOS_OK) throw RuntimeError("bad redshift");
        return dc;
    }
    double getTransverseComovingScale(double z) const
    {
        double dm = 0;
        if (tos_cosmo_distances(ol_, om_, 1, &z, nullptr, &dm) != TOS_OK) throw RuntimeError("bad redshift");
        return dm;
    }
    /// The distance transform of h5healxi.cc:196-216 on HIP device `device`:
    /// distance[i] = getLineOfSightComovingDistance(10^(loglam[i] - logLyA) - 1)
    /// rounded to float, for n pixels (added to the reference's interface).
    void transformPixels(int64_t n, const float *loglam, float *distance, int device = 0) const
    {
        tos_ctx *ctx = nullptr;
        if (tos_ctx_create(device, &ctx) != TOS_OK) throw RuntimeError("LambdaCdmUniverse: no usable HIP device");
        const int rc = tos_pixel_distances(ctx, ol_, om_, n, loglam, distance);
        const std::string err = rc == TOS_OK ? std::string() : std::string(tos_last_error(ctx));
        tos_ctx_destroy(ctx);
        if (rc != TOS_OK) throw RuntimeError("LambdaCdmUniverse: " + err);
    }
private:
    double ol_, om_;
};

}  // namespace turbooctospice

#endif
