/*
 * healpix_oracle.c -- HEALPix RING-scheme pixelisation used by HealpixBins
 * (turbooctospice/HealpixBins.h:23-72; healpix_cxx 3.30, not vendored).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the published RING algorithm (Gorski et al. 2005, ApJ 622, 759,
 * section 4 and appendix): equatorial belt |z| <= 2/3 indexed by the two
 * families of edge lines, polar caps by the distance to the pole.
 * Pinning: pix2ang(ang2pix) round trip for every pixel centre at
 * nside 1, 2, 4, 16, 64; pixel 0 and npix-1 at the poles; ring starts on
 * the equator; equal-area occupancy of a uniform sample.  The exact cell
 * chosen for points lying on a cell boundary is unpinned (healpix_cxx's
 * floating-point details are not available here); it only moves a
 * sightline between sky-bin subsamples, never changes the xi totals.
 */
#include "oracle.h"

#include <math.h>

static const double ORA_PI = 3.141592653589793238462643383279502884;

int64_t ora_ang2pix_ring(int64_t nside, double theta, double phi)
{
    const double z = cos(theta);
    const double za = fabs(z);
    double tt = fmod(phi, 2.0 * ORA_PI);
    if (tt < 0) tt += 2.0 * ORA_PI;
    tt = tt / (0.5 * ORA_PI);                 /* in [0, 4) */
    if (tt >= 4.0) tt = 0.0;
    const int64_t npix = 12 * nside * nside;
    const int64_t ncap = 2 * nside * (nside - 1);
    if (za <= 2.0 / 3.0) {
        const int64_t nl4 = 4 * nside;
        const double a = (double)nside * (0.5 + tt);
        const double b = (double)nside * z * 0.75;
        const int64_t jp = (int64_t)(a - b);  /* ascending edge line */
        const int64_t jm = (int64_t)(a + b);  /* descending edge line */
        const int64_t ir = nside + 1 + jp - jm;       /* ring in [1, 2 nside + 1] */
        const int64_t kshift = 1 - (ir & 1);
        int64_t ip = (jp + jm - nside + kshift + 1) / 2;
        ip = ((ip % nl4) + nl4) % nl4;
        return ncap + (ir - 1) * nl4 + ip;
    }
    const double tp = tt - floor(tt);
    const double tmp = (double)nside * sqrt(3.0 * (1.0 - za));
    const int64_t jp = (int64_t)(tp * tmp);
    const int64_t jm = (int64_t)((1.0 - tp) * tmp);
    const int64_t ir = jp + jm + 1;           /* ring counted from the nearest pole */
    int64_t ip = (int64_t)(tt * (double)ir);
    if (ip >= 4 * ir) ip -= 4 * ir;
    if (z > 0) return 2 * ir * (ir - 1) + ip;
    return npix - 2 * ir * (ir + 1) + ip;
}

static int64_t isqrt64(int64_t v)
{
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

void ora_pix2ang_ring(int64_t nside, int64_t pix, double *theta, double *phi)
{
    const int64_t npix = 12 * nside * nside;
    const int64_t ncap = 2 * nside * (nside - 1);
    const double fn = (double)nside;
    double z, ph;
    if (pix < ncap) {
        const int64_t ir = (1 + isqrt64(1 + 2 * pix)) / 2;
        const int64_t iphi = pix + 1 - 2 * ir * (ir - 1);
        z = 1.0 - (double)(ir * ir) / (3.0 * fn * fn);
        ph = ((double)iphi - 0.5) * ORA_PI / (2.0 * (double)ir);
    } else if (pix < npix - ncap) {
        const int64_t ip = pix - ncap;
        const int64_t ir = ip / (4 * nside) + nside;
        const int64_t iphi = ip % (4 * nside) + 1;
        const double fodd = ((ir + nside) & 1) ? 1.0 : 0.5;
        z = (double)(2 * nside - ir) * 2.0 / (3.0 * fn);
        ph = ((double)iphi - fodd) * ORA_PI / (2.0 * fn);
    } else {
        const int64_t ip = npix - pix;
        const int64_t ir = (1 + isqrt64(2 * ip - 1)) / 2;
        const int64_t iphi = 4 * ir + 1 - (ip - 2 * ir * (ir - 1));
        z = -1.0 + (double)(ir * ir) / (3.0 * fn * fn);
        ph = ((double)iphi - 0.5) * ORA_PI / (2.0 * (double)ir);
    }
    *theta = acos(z);
    *phi = ph;
}

int64_t ora_radec2pix(int64_t nside, double ra, double dec)
{
    /* radec2pnt: theta = pi/2 - dec, phi = ra (HealpixBins.h:69-71) */
    const double pi = atan(1.0) * 4;   /* constants.h:12 */
    return ora_ang2pix_ring(nside, 0.5 * pi - dec, ra);
}
