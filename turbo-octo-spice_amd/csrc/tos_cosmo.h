// tos_cosmo.h -- the FP64 distance transform of h5healxi.cc:196-216,
// shared by the host and the device.
//
// D_C(z) = (c/H0) int_0^z dz'/E(z'), E^2 = Om (1+z)^3 + Ok (1+z)^2 + OL
// (cosmo::LambdaCdmUniverse, h5healxi.cc:123-131), integrated with composite
// 8-point Gauss-Legendre in ln(1+z) over 64 panels.  Everything here uses
// only IEEE-rounded operations (+, -, *, /, sqrt) and exact ones (floor,
// frexp, ldexp) in a fixed order -- exp and log1p are written out below
// rather than taken from the host's or the device's math library -- so the
// host and the device compute the same bits (the library is built with
// -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace tos {

#define TOS_HD __host__ __device__ inline

// exp(x): x = k ln2 + r (Cody-Waite, |r| <= ln2 / 2), exp(r) by its Taylor
// series to r^14 (0.35^15 / 15! < 1e-20), scaled by 2^k; NaN stays NaN, x
// below -745 gives 0 and x above 709 infinity (the transform's arguments
// lie in [0, 2]).
TOS_HD double exp_rn(double x)
{
    if (!(x > -745.0)) return x < 0 ? 0.0 : x;
    if (x > 709.0) return HUGE_VAL;
    const double ln2_hi = 6.93147180369123816490e-01;   // 32 trailing zero bits: k ln2_hi is exact
    const double ln2_lo = 1.90821492927058770002e-10;
    const double k = floor(x * 1.44269504088896338700e+00 + 0.5);
    const double r = (x - k * ln2_hi) - k * ln2_lo;
    double p = 1.0 / 87178291200.0;   // 1/14!
    const double inv_fact[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                 1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,
                                 1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,        0.5,
                                 1.0,                1.0};
    for (int j = 0; j < 14; ++j) p = p * r + inv_fact[j];
    return ldexp(p, (int)k);
}

// log(1 + z) for z > -1: 1 + z = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| < 0.1716, by its series to
// s^29; the rounding of 1 + z is corrected by (z - (u - 1)) / u.
TOS_HD double log1p_rn(double z)
{
    const double u = 1.0 + z;
    if (u == 1.0) return z;
    int e;
    double m = frexp(u, &e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double c = (z - (u - 1.0)) / u;
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double s2 = s * s;
    double q = 1.0 / 29.0;
    for (int j = 27; j >= 1; j -= 2) q = q * s2 + 1.0 / (double)j;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return ((double)e * ln2_lo + c + 2.0 * s * q) + (double)e * ln2_hi;
}

TOS_HD double hubble_e(double ol, double om, double z)
{
    const double ok = 1.0 - om - ol;
    const double a = 1.0 + z;
    return sqrt(om * a * a * a + ok * a * a + ol);
}

// D_C(z) in Mpc/h.
TOS_HD double los_comoving(double ol, double om, double z)
{
    const double xg[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290, -0.1834346424956498,
                          0.1834346424956498,  0.5255324099163290,  0.7966664774136267,  0.9602898564975363};
    const double wg[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620,
                          0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
    const double c_over_h0 = 2997.92458;
    if (z == 0) return 0.0;
    const double L = log1p_rn(z);
    const int panels = 64;
    const double h = L / panels;
    double acc = 0;
    for (int p = 0; p < panels; ++p) {
        const double mid = (p + 0.5) * h;
        for (int k = 0; k < 8; ++k) {
            const double x = mid + 0.5 * h * xg[k];
            const double a = exp_rn(x);   // 1 + z'
            acc += wg[k] * a / hubble_e(ol, om, a - 1.0);
        }
    }
    return c_over_h0 * acc * 0.5 * h;
}

// The pixel distance of h5healxi.cc:200-203: D_C(10^(loglam - log10 Lya) - 1)
// (loglamToRedshift, h5healxi.cc:38-40), rounded to the pixel's float.
TOS_HD float pixel_distance(double ol, double om, double log_lya, float loglam)
{
    const double z = exp_rn(((double)loglam - log_lya) * 2.30258509299404568402) - 1.0;
    return (float)los_comoving(ol, om, z);
}

#undef TOS_HD

}  // namespace tos
