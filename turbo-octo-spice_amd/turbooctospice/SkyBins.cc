// SkyBins.cc -- HEALPix and plate sky bins of the turbooctospice C++ API.
#include "SkyBins.h"

#include <fitsio.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>

namespace turbooctospice {

namespace {
struct V3 {
    double x, y, z;
};
V3 unit(double ra, double dec) { return V3{std::cos(dec) * std::cos(ra), std::cos(dec) * std::sin(ra), std::sin(dec)}; }
double angle(V3 a, V3 b)
{
    const double cx = a.y * b.z - a.z * b.y, cy = a.z * b.x - a.x * b.z, cz = a.x * b.y - a.y * b.x;
    return std::atan2(std::sqrt(cx * cx + cy * cy + cz * cz), a.x * b.x + a.y * b.y + a.z * b.z);
}
V3 cell_centre(long nside, int pix);
}  // namespace

// pix2ang in the RING scheme is exported by the device library's host part
extern "C" int tos_healpix_radec2pix(int32_t order, int64_t n, const double *ra, const double *dec, int32_t *out_pix);

namespace {
// RING pixel centre (Gorski et al. 2005, eqs. 8-9)
V3 cell_centre(long nside, int pix)
{
    const long npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    const double fn = (double)nside;
    double z, ph;
    if (pix < ncap) {
        long ring = (long)((1 + std::sqrt(1.0 + 2.0 * pix)) / 2);
        while (2 * ring * (ring - 1) > pix) --ring;
        while (2 * (ring + 1) * ring <= pix) ++ring;
        const long k = pix + 1 - 2 * ring * (ring - 1);
        z = 1.0 - (double)(ring * ring) / (3.0 * fn * fn);
        ph = ((double)k - 0.5) * pi / (2.0 * (double)ring);
    } else if (pix < npix - ncap) {
        const long p = pix - ncap, ring = p / (4 * nside) + nside, k = p % (4 * nside) + 1;
        const double shift = ((ring + nside) & 1) ? 1.0 : 0.5;
        z = (double)(2 * nside - ring) * 2.0 / (3.0 * fn);
        ph = ((double)k - shift) * pi / (2.0 * fn);
    } else {
        const long p = npix - pix;
        long ring = (long)((1 + std::sqrt(2.0 * p - 1)) / 2);
        while (2 * ring * (ring - 1) >= p) --ring;
        while (2 * (ring + 1) * ring < p) ++ring;
        const long k = 4 * ring + 1 - (p - 2 * ring * (ring - 1));
        z = -1.0 + (double)(ring * ring) / (3.0 * fn * fn);
        ph = ((double)k - 0.5) * pi / (2.0 * (double)ring);
    }
    const double st = std::sqrt(std::max(0.0, 1.0 - z * z));
    return V3{st * std::cos(ph), st * std::sin(ph), z};
}
}  // namespace

HealpixBinsI::HealpixBinsI(int order) : SkyBins<int>(), order_(order), nside_(1L << order)
{
    if (order < 0 || order > 13) throw RuntimeError("HealpixBins: order must be in [0, 13]");
}

int HealpixBinsI::ang2pix(double ra, double dec) const
{
    int32_t pix = 0;
    tos_healpix_radec2pix(order_, 1, &ra, &dec, &pix);
    return pix;
}

void HealpixBinsI::addItem(const Forest &forest, const int &item)
{
    const int key = ang2pix(forest.ra, forest.dec);
    insert(key, item);
    double &r = cell_radius_[key];
    r = std::max(r, angle(cell_centre(nside_, key), unit(forest.ra, forest.dec)));
}

std::vector<int> HealpixBinsI::getBinIndicesWithinRadius(double ra, double dec, double radius) const
{
    // every occupied cell that can hold a member within `radius`, ascending
    std::vector<int> out;
    const V3 p = unit(ra, dec);
    for (const auto &kv : cell_radius_)
        if (angle(p, cell_centre(nside_, kv.first)) <= radius + kv.second + 1e-9) out.push_back(kv.first);
    return out;
}

void HealpixBinsI::printBinCenter(const int index) const
{
    const V3 c = cell_centre(nside_, index);
    std::cout << std::asin(c.z) << " " << std::atan2(c.y, c.x);
}

PlateBinsI::PlateBinsI(PlateList const &plates) : SkyBins<int>(), platelist_(plates) {}

PlateBinsI::PlateBinsI(std::string const &filename) : SkyBins<int>()
{
    fitsfile *f = nullptr;
    int status = 0;
    if (fits_open_table(&f, filename.c_str(), READONLY, &status) == 0) {
        long nrows = 0;
        int cp = 0, cr = 0, cd = 0;
        fits_get_num_rows(f, &nrows, &status);
        fits_get_colnum(f, CASEINSEN, (char *)"PLATE", &cp, &status);
        fits_get_colnum(f, CASEINSEN, (char *)"RACEN", &cr, &status);
        fits_get_colnum(f, CASEINSEN, (char *)"DECCEN", &cd, &status);
        std::vector<int> plate((size_t)nrows);
        std::vector<float> rac((size_t)nrows), decc((size_t)nrows);
        fits_read_col(f, TINT, cp, 1, 1, nrows, nullptr, plate.data(), nullptr, &status);
        fits_read_col(f, TFLOAT, cr, 1, 1, nrows, nullptr, rac.data(), nullptr, &status);
        fits_read_col(f, TFLOAT, cd, 1, 1, nrows, nullptr, decc.data(), nullptr, &status);
        int st2 = 0;
        fits_close_file(f, &st2);
        if (status) throw RuntimeError("PlateBins: cannot read PLATE/RACEN/DECCEN from " + filename);
        // centres read as float degrees (PlateBins.h:38-45)
        for (long i = 0; i < nrows; ++i)
            platelist_[plate[(size_t)i]] = SkyObject(rac[(size_t)i] * deg2rad, decc[(size_t)i] * deg2rad);
        return;
    }
    std::ifstream in(filename);
    if (!in) throw RuntimeError("PlateBins: cannot open " + filename);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        int p;
        float ra, dec;
        if (ss >> p >> ra >> dec) platelist_[p] = SkyObject(ra * deg2rad, dec * deg2rad);
    }
}

void PlateBinsI::addItem(const Forest &forest, const int &item) { insert(forest.plate, item); }

std::vector<int> PlateBinsI::getBinIndicesWithinRadius(double ra, double dec, double radius) const
{
    std::vector<int> out;
    const SkyObject p(ra, dec);
    for (const auto &plate : platelist_) {
        const double sep = std::acos(p.angularSeparation(plate.second));
        if (std::fabs(sep - 1.5 * deg2rad) < radius) out.push_back(plate.first);
    }
    return out;
}

void PlateBinsI::printBinCenter(const int index) const
{
    const SkyObject &p = platelist_.at(index);
    std::cout << p.dec << " " << p.ra << std::endl;
}

}  // namespace turbooctospice
