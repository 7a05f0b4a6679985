// types.h -- value types of the turbooctospice C++ API (mirrors
// turbooctospice/types.h:15-127 of the reference: same names, fields and
// semantics; std::shared_ptr replaces boost::shared_ptr).
#ifndef TOS_AMD_TURBOOCTOSPICE_TYPES
#define TOS_AMD_TURBOOCTOSPICE_TYPES

#include <cmath>
#include <memory>
#include <vector>

#include "RuntimeError.h"

namespace turbooctospice {

/// An object on the sky; ra and dec in radians (types.h:15-40).
struct SkyObject {
    double ra = 0, dec = 0, sin_dec = 0, cos_dec = 0, sin_ra = 0, cos_ra = 0;
    SkyObject() = default;
    SkyObject(double ra_, double dec_) : ra(ra_), dec(dec_)
    {
        sin_dec = std::sin(dec);
        cos_dec = std::cos(dec);
        sin_ra = std::sin(ra);
        cos_ra = std::cos(ra);
    }
    /// cos of the angular separation, evaluated in the reference's order.
    double angularSeparation(SkyObject const &other) const
    {
        const double t1 = sin_dec * other.sin_dec;
        const double t2 = cos_dec * other.cos_dec;
        const double t3 = sin_ra * other.sin_ra;
        const double t4 = cos_ra * other.cos_ra;
        return t1 + t2 * (t3 + t4);
    }
};

/// A forest pixel (types.h:43-52).
struct ForestPixel {
    float value, loglam, weight, distance;
    ForestPixel(float val, float loglambda, float wgt, float dist)
        : value(val), loglam(loglambda), weight(wgt), distance(dist)
    {
    }
};

/// A forest sightline (types.h:55-66).
struct Forest : SkyObject {
    std::vector<ForestPixel> pixels;
    int forest_id, plate;
    Forest(double ra_, double dec_, int forest_id_, int plate_)
        : SkyObject(ra_, dec_), forest_id(forest_id_), plate(plate_)
    {
    }
};

/// A correlation-function bin (types.h:69-121).
struct XiBin {
    double didj = 0, di = 0, dj = 0, wgt = 0, wi = 0, wj = 0, z = 0;
    long num_pairs = 0;
    void accumulate_pair(ForestPixel const &i, ForestPixel const &j)
    {
        const float wdi = i.weight * i.value, wdj = j.weight * j.value;
        const float p = wdi * wdj, ww = i.weight * j.weight, ls = i.loglam + j.loglam;
        didj += p;
        di += wdi;
        dj += wdj;
        wgt += ww;
        wi += i.weight;
        wj += j.weight;
        z += 0.5 * (double)ls;
        ++num_pairs;
    }
    double getMeanProduct()
    {
        if (wgt <= 0) throw RuntimeError("XiBin: wgt <= 0");
        return didj / wgt;
    }
    void finalize()
    {
        didj = wgt > 0 ? didj / wgt : 0;
        di = wi > 0 ? di / wi : 0;
        dj = wj > 0 ? dj / wj : 0;
        z = num_pairs > 0 ? z / num_pairs : 0;
    }
    XiBin &operator+=(const XiBin &rhs)
    {
        didj += rhs.didj;
        di += rhs.di;
        dj += rhs.dj;
        wgt += rhs.wgt;
        wi += rhs.wi;
        wj += rhs.wj;
        z += rhs.z;
        num_pairs += rhs.num_pairs;
        return *this;
    }
    friend XiBin operator+(XiBin lhs, const XiBin &rhs) { return lhs += rhs; }
};

class AbsTwoPointGrid;
typedef std::shared_ptr<AbsTwoPointGrid> AbsTwoPointGridPtr;

template <typename T> class SkyBins;
typedef std::shared_ptr<SkyBins<int>> SkyBinsIPtr;

}  // namespace turbooctospice

#endif
