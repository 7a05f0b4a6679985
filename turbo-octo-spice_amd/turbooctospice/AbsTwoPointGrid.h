// AbsTwoPointGrid.h -- binning grids of the estimator (mirrors the
// reference's AbsTwoPointGrid.h:17-58, CartesianGrid.h, PolarGrid.h,
// QuasarGrid.h).  likely::createBinning / likely::BinnedGrid are not
// vendored by the reference, so the uniform "[min:max]*n" binning they
// provide is implemented here (UniformBinning, createBinning).
#ifndef TOS_AMD_TURBOOCTOSPICE_ABS_TWO_POINT_GRID
#define TOS_AMD_TURBOOCTOSPICE_ABS_TWO_POINT_GRID

#include <memory>
#include <string>
#include <vector>

#include "constants.h"
#include "tos_amd.h"
#include "types.h"

namespace turbooctospice {

/// A uniform binning of [min, max) into n bins.
class UniformBinning {
public:
    UniformBinning(int nbins, double min, double max);
    int getNBins() const { return n_; }
    double getBinLowEdge(int i) const { return min_ + i * width_; }
    double getBinHighEdge(int i) const { return min_ + (i + 1) * width_; }
    double getBinWidth(int i = 0) const { (void)i; return width_; }
    double getBinCenter(int i) const { return min_ + (i + 0.5) * width_; }
private:
    int n_;
    double min_, width_;
};
typedef std::shared_ptr<const UniformBinning> AbsBinningCPtr;

/// Parses "[min:max]*n" (likely::createBinning's uniform form).
AbsBinningCPtr createBinning(std::string const &spec);

/// A 3-axis grid for two-point functions (AbsTwoPointGrid.h:17-58).
class AbsTwoPointGrid {
public:
    AbsTwoPointGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3);
    virtual ~AbsTwoPointGrid();
    int getNBinsTotal() const;
    void getBinCenters(int index, std::vector<double> &binCenters) const;
    virtual double minAngularScale(double scale) const = 0;
    virtual double maxAngularScale(double scale) const = 0;
    /// Separation of a pixel pair along each axis, with the grid's range
    /// checks (CartesianGrid.cc:16-27, PolarGrid.cc:16-26, QuasarGrid.cc:14-23).
    virtual bool getSeparation(ForestPixel const &a, ForestPixel const &b, double const &cosij,
                               double const &thetaij, std::vector<double> &separation) const = 0;
    /// Flat bin index of a pixel pair with the estimator's arithmetic
    /// (XiEstimator.cc:168-294).  The reference leaves this a stub; here it
    /// is the host-side counterpart of the device kernel.
    virtual bool getBinIndex(ForestPixel const &a, ForestPixel const &b, double const &cosij,
                             double const &thetaij, int &binIndex) const;
    int getAxisNBins(int axis) const;
    double getAxisMax(int axis) const;
    double getAxisMin(int axis) const;
    double getAxisBinWidth(int axis, int index = 0) const;
    /// The grid as passed to the device library.
    virtual int coordinateType() const = 0;
    tos_grid toC() const;
protected:
    std::vector<double> xmin, xmax;
    std::vector<AbsBinningCPtr> axes_;
};

/// (r_par, r_perp, log10(lya(1+z))) grid (CartesianGrid.h:10-46).
class CartesianGrid : public AbsTwoPointGrid {
public:
    enum AxisLabels { RPara, RPerp, LogLya1pz };
    CartesianGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3);
    double minAngularScale(double scale) const override { return xmin[RPerp] / scale; }
    double maxAngularScale(double scale) const override { return xmax[RPerp] / scale; }
    bool getSeparation(ForestPixel const &a, ForestPixel const &b, double const &cosij, double const &thetaij,
                       std::vector<double> &separation) const override;
    int coordinateType() const override { return TOS_CARTESIAN; }
private:
    double x1minSq, x1maxSq;
};

/// (r, mu, log10(lya(1+z))) grid (PolarGrid.h:10-46).
class PolarGrid : public AbsTwoPointGrid {
public:
    enum AxisLabels { R, Mu, LogLya1pz };
    PolarGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3);
    double minAngularScale(double scale) const override
    {
        return xmin[R] * std::sqrt(1.0 - xmax[Mu] * xmax[Mu]) / scale;
    }
    double maxAngularScale(double scale) const override
    {
        return xmax[R] * std::sqrt(1.0 - xmin[Mu] * xmin[Mu]) / scale;
    }
    bool getSeparation(ForestPixel const &a, ForestPixel const &b, double const &cosij, double const &thetaij,
                       std::vector<double> &separation) const override;
    int coordinateType() const override { return TOS_POLAR; }
private:
    double x1minSq, x1maxSq;
};

/// (dloglam, theta, log10(lya(1+z))) grid in observing coordinates
/// (QuasarGrid.h:10-45); the theta axis is in radians.
class QuasarGrid : public AbsTwoPointGrid {
public:
    enum AxisLabels { DLogLam, Theta, LogLya1pz };
    QuasarGrid(AbsBinningCPtr axis1, AbsBinningCPtr axis2, AbsBinningCPtr axis3);
    double minAngularScale(double) const override { return xmin[Theta]; }
    double maxAngularScale(double) const override { return xmax[Theta]; }
    bool getSeparation(ForestPixel const &a, ForestPixel const &b, double const &cosij, double const &thetaij,
                       std::vector<double> &separation) const override;
    int coordinateType() const override { return TOS_OBSERVING; }
};

}  // namespace turbooctospice

#endif
