// AbsTwoPointGrid.cc -- grids of the turbooctospice C++ API.
#include "AbsTwoPointGrid.h"

#include <cstdio>
#include <cstring>

namespace turbooctospice {

UniformBinning::UniformBinning(int nbins, double min, double max) : n_(nbins), min_(min)
{
    if (nbins <= 0 || !(max > min)) throw RuntimeError("UniformBinning: need nbins > 0 and max > min");
    width_ = (max - min) / nbins;
}

AbsBinningCPtr createBinning(std::string const &spec)
{
    tos_grid g;
    std::memset(&g, 0, sizeof(g));
    if (tos_grid_parse_axis(spec.c_str(), &g, 0) != TOS_OK)
        throw RuntimeError("createBinning: expected \"[min:max]*n\", got \"" + spec + "\"");
    double lo, hi;
    int n;
    std::sscanf(spec.c_str(), " [ %lf : %lf ] * %d", &lo, &hi, &n);
    return std::make_shared<UniformBinning>(n, lo, hi);
}

AbsTwoPointGrid::AbsTwoPointGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3)
    : axes_{axis1, axis2, axis3}
{
    for (int axis = 0; axis < 3; ++axis) {
        if (!axes_[axis]) throw RuntimeError("AbsTwoPointGrid: missing axis binning");
        xmin.push_back(axes_[axis]->getBinLowEdge(0));
        xmax.push_back(axes_[axis]->getBinHighEdge(axes_[axis]->getNBins() - 1));
    }
}

AbsTwoPointGrid::~AbsTwoPointGrid() {}

int AbsTwoPointGrid::getNBinsTotal() const
{
    return axes_[0]->getNBins() * axes_[1]->getNBins() * axes_[2]->getNBins();
}

void AbsTwoPointGrid::getBinCenters(int index, std::vector<double> &binCenters) const
{
    // row-major, axis 0 slowest (XiEstimator.cc:249,254)
    binCenters.resize(3);
    const int n1 = axes_[1]->getNBins(), n2 = axes_[2]->getNBins();
    const int i2 = index % n2, i1 = (index / n2) % n1, i0 = index / (n1 * n2);
    binCenters[0] = axes_[0]->getBinCenter(i0);
    binCenters[1] = axes_[1]->getBinCenter(i1);
    binCenters[2] = axes_[2]->getBinCenter(i2);
}

int AbsTwoPointGrid::getAxisNBins(int axis) const { return axes_.at(axis)->getNBins(); }
double AbsTwoPointGrid::getAxisMin(int axis) const { return axes_.at(axis)->getBinLowEdge(0); }
double AbsTwoPointGrid::getAxisMax(int axis) const
{
    return axes_.at(axis)->getBinHighEdge(getAxisNBins(axis) - 1);
}
double AbsTwoPointGrid::getAxisBinWidth(int axis, int index) const { return axes_.at(axis)->getBinWidth(index); }

tos_grid AbsTwoPointGrid::toC() const
{
    tos_grid g;
    std::memset(&g, 0, sizeof(g));
    g.type = coordinateType();
    for (int k = 0; k < 3; ++k) {
        g.nbins[k] = getAxisNBins(k);
        g.min[k] = getAxisMin(k);
        g.width[k] = getAxisBinWidth(k);
        g.max[k] = getAxisMax(k);
    }
    return g;
}

namespace {
// x86-64 static_cast<unsigned>(double): cvttsd2si r64, low 32 bits
unsigned cast_unsigned(double x)
{
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) return 0u;
    return (unsigned)(unsigned long long)(long long)x;
}
}  // namespace

bool AbsTwoPointGrid::getBinIndex(ForestPixel const &a, ForestPixel const &b, double const &cosij,
                                  double const &thetaij, int &binIndex) const
{
    const unsigned n0 = getAxisNBins(0), n1 = getAxisNBins(1), n2 = getAxisNBins(2);
    const double inv0 = 1.0 / getAxisBinWidth(0), inv1 = 1.0 / getAxisBinWidth(1);
    const double inv_z = 1.0 / (double)n2;   // XiEstimator.cc:232 uses 1/NBins
    const double zmin = getAxisMin(2);
    const float d1sq = a.distance * a.distance;
    const double r1_sq = d1sq;
    const float two_d1 = 2.0f * a.distance;
    const double two_r1_cos12 = (double)two_d1 * cosij;
    unsigned bin;
    const int type = coordinateType();
    if (type == TOS_POLAR) {
        const double d2 = b.distance;
        const double r = std::sqrt(r1_sq + (d2 - two_r1_cos12) * d2);
        bin = cast_unsigned((r - getAxisMin(0)) * inv0);
        if (bin >= n0 || r == 0) return false;
        const float adiff = std::fabs(a.distance - b.distance);
        const unsigned imu = cast_unsigned(((double)adiff / r - getAxisMin(1)) * inv1);
        if (imu >= n1) return false;
        bin = imu + bin * n1;
    } else if (type == TOS_CARTESIAN) {
        const float adiff = std::fabs(a.distance - b.distance);
        const double rpar = adiff;
        bin = cast_unsigned((rpar - getAxisMin(0)) * inv0);
        if (bin >= n0) return false;
        const double d2 = b.distance;
        const double r_sq = r1_sq + (d2 - two_r1_cos12) * d2;
        const unsigned perp = cast_unsigned((std::sqrt(std::fabs(r_sq - rpar * rpar)) - getAxisMin(1)) * inv1);
        if (perp >= n1) return false;
        bin = perp + bin * n1;
    } else {
        const unsigned ith = cast_unsigned((thetaij - getAxisMin(1)) * inv1);
        const float adl = std::fabs(a.loglam - b.loglam);
        bin = cast_unsigned(((double)adl - getAxisMin(0)) * inv0);
        if (bin >= n0) return false;
        bin = ith + bin * n1;
    }
    const float lsum = a.loglam + b.loglam;
    const unsigned iz = cast_unsigned((0.5 * (double)lsum - zmin) * inv_z);
    bin = iz + bin * n2;
    if (bin >= (unsigned)getNBinsTotal()) return false;
    binIndex = (int)bin;
    return true;
}

CartesianGrid::CartesianGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3)
    : AbsTwoPointGrid(axis1, axis2, axis3)
{
    x1minSq = xmin[0] * xmin[0];
    x1maxSq = xmax[0] * xmax[0];
}

bool CartesianGrid::getSeparation(ForestPixel const &a, ForestPixel const &b, double const &cosij, double const &,
                                  std::vector<double> &separation) const
{
    // CartesianGrid.cc:16-27
    separation.resize(3);
    const double distSq = a.distance * a.distance + b.distance * b.distance - 2 * a.distance * b.distance * cosij;
    if (distSq >= x1maxSq || distSq < x1minSq) return false;
    separation[0] = std::fabs(a.distance - b.distance);
    separation[1] = std::sqrt(distSq - separation[0] * separation[0]);
    if (separation[1] < xmin[1] || separation[1] >= xmax[1]) return false;
    separation[2] = 0.5 * (a.loglam + b.loglam) - logLyA;
    if (separation[2] < xmin[2] || separation[2] >= xmax[2]) return false;
    return true;
}

PolarGrid::PolarGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3)
    : AbsTwoPointGrid(axis1, axis2, axis3)
{
    x1minSq = xmin[0] * xmin[0];
    x1maxSq = xmax[0] * xmax[0];
}

bool PolarGrid::getSeparation(ForestPixel const &a, ForestPixel const &b, double const &cosij, double const &,
                              std::vector<double> &separation) const
{
    // PolarGrid.cc:16-26
    separation.resize(3);
    const double distSq = a.distance * a.distance + b.distance * b.distance - 2 * a.distance * b.distance * cosij;
    if (distSq >= x1maxSq || distSq < x1minSq) return false;
    separation[0] = std::sqrt(distSq);
    separation[1] = std::fabs(a.distance - b.distance) / separation[0];
    if (separation[1] < xmin[1] || separation[1] >= xmax[1]) return false;
    separation[2] = 0.5 * (a.loglam + b.loglam) - logLyA;
    if (separation[2] < xmin[2] || separation[2] >= xmax[2]) return false;
    return true;
}

QuasarGrid::QuasarGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3)
    : AbsTwoPointGrid(axis1, axis2, axis3)
{
}

bool QuasarGrid::getSeparation(ForestPixel const &a, ForestPixel const &b, double const &, double const &thetaij,
                               std::vector<double> &separation) const
{
    // QuasarGrid.cc:14-23 (theta in arcmin there; the estimator uses radians)
    separation.resize(3);
    separation[0] = std::fabs(a.loglam - b.loglam);
    if (separation[0] >= xmax[0] || separation[0] < xmin[0]) return false;
    separation[1] = thetaij * rad2arcmin;
    separation[2] = 0.5 * (a.loglam + b.loglam) - logLyA;
    if (separation[2] < xmin[2] || separation[2] >= xmax[2]) return false;
    return true;
}

}  // namespace turbooctospice
