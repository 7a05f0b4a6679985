// Cosmology.h -- the LCDM distances used by the h5healxi distance transform
// (h5healxi.cc:123-131, 196-239; cosmo::LambdaCdmUniverse is not vendored
// by the reference).  Distances in Mpc/h.
#ifndef TOS_AMD_TURBOOCTOSPICE_COSMOLOGY
#define TOS_AMD_TURBOOCTOSPICE_COSMOLOGY

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
private:
    double ol_, om_;
};

}  // namespace turbooctospice

#endif
