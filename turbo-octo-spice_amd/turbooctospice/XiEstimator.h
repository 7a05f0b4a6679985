// XiEstimator.h -- the pair-counting correlation estimator (mirrors the
// reference's XiEstimator.h:17-86: same constructor, run, save_results,
// save_subsamples and print_stats).  The pixel-pair pass runs on the GPU
// through the C-ABI of tos_amd.h; the candidate sightline pairs come from
// the SkyBins object exactly as in XiEstimator::skybin_xi_task.
#ifndef TOS_AMD_TURBOOCTOSPICE_XI_ESTIMATOR
#define TOS_AMD_TURBOOCTOSPICE_XI_ESTIMATOR

#include <map>
#include <string>
#include <vector>

#include "AbsTwoPointGrid.h"
#include "SkyBins.h"
#include "constants.h"
#include "tos_amd.h"
#include "types.h"

namespace turbooctospice {

class XiEstimator {
public:
    enum BinningCoordinateType { PolarCoordinates, CartesianCoordinates, ObservingCoordinates };
    /// Arithmetic of the pixel-pair pass (an addition to the reference API).
    enum Precision { FP64 = TOS_FP64, FP32C = TOS_FP32C };

    XiEstimator(double scale, AbsTwoPointGridPtr grid, BinningCoordinateType type, std::vector<Forest> sightlines,
                SkyBinsIPtr skybins);
    ~XiEstimator();
    XiEstimator(const XiEstimator &) = delete;
    XiEstimator &operator=(const XiEstimator &) = delete;

    /// Runs the estimator.  nthreads host threads enumerate candidate pairs;
    /// the pixel pairs run on the GPU.  do_cov computes the covariance of the
    /// sky-bin subsamples.
    void run(int nthreads, bool do_cov = false);
    void save_results(std::string outfile) const;
    void save_subsamples(std::string outfile_base) const;
    void print_stats() const;

    // ---- additions to the reference API ----
    void setPrecision(Precision p) { precision_ = p; }
    /// Enumerate the candidate pairs on the GPU (tos_pairs_build) when the
    /// sky bins are HealpixBins or PlateBins.  The kept pairs, their cos
    /// values and every histogram are those of the host enumeration; the
    /// "considered" count is the device neighbour superset's.
    void setDeviceSearch(bool on) { device_search_ = on; }
    void setDevice(int device) { device_ = device; }
    /// Keep per-sky-bin subsamples (needed by save_subsamples and do_cov).
    void setKeepSubsamples(bool keep) { keep_subsamples_ = keep; }
    const std::vector<XiBin> &getXi() const { return xi_; }
    const std::map<int, std::vector<XiBin>> &getSubsamples() const { return skybin_xis_; }
    const std::vector<double> &getCovariance() const { return cov_; }
    unsigned long numPixelPairs() const { return num_pixel_pairs_; }
    unsigned long numPixelPairsUsed() const { return num_pixel_pairs_used_; }
    unsigned long numSightlinePairsUsed() const { return num_sightline_pairs_used_; }
    double maxAngularScale() const { return max_ang_; }

private:
    struct Csr {
        std::vector<int32_t> row_los, row_skybin, other;
        std::vector<int64_t> row_ptr;
        std::vector<double> csep;
    };
    /// Host enumeration through the SkyBins object (XiEstimator.cc:121-159).
    void enumerate_host(int nthreads, Csr &csr);
    /// The pixel-pair passes; csr == nullptr: the candidates come from the
    /// GPU search over the sky bins.
    void accumulate(const Csr *csr, int nthreads);
    unsigned long num_sightline_pairs_, num_sightline_pairs_used_, num_pixels_, num_pixel_pairs_,
        num_pixel_pairs_used_;
    double max_ang_, cos_max_ang_;
    const BinningCoordinateType coordinate_type_;
    const AbsTwoPointGridPtr grid_;
    const SkyBinsIPtr skybins_;
    const std::vector<Forest> sightlines_;
    const unsigned num_xi_bins_;
    std::map<int, std::vector<XiBin>> skybin_xis_;
    std::vector<XiBin> xi_;
    std::vector<double> cov_;
    bool cov_good_;
    Precision precision_ = FP64;
    int device_ = 0;
    bool keep_subsamples_ = true;
    bool device_search_ = false;
};

/// Binning arithmetic of bruteGPU: the reference kernel's FP32 positions
/// and FP32 bin spacing with under/overflow slots (gpu/brute.cu:26-55), or
/// the FP64 pair binning of old/BinPair.h:28-47.
enum BruteArithmetic { BruteFP32 = 0, BruteFP64 = 1 };

/// The GPU 3D brute-force estimator (gpu/brute.h:20): xi per bin of
/// columns {x, y, z, d, w}, xi[b] = sum w_i w_j d_i d_j / sum w_i w_j over
/// the pairs i < j in bin b (gpu/brute.cu:184-212).  The reference's FP32
/// binning by default.  Unlike the reference the input columns are not
/// padded; chunksize is accepted for signature compatibility and ignored
/// (all points stay resident on the device).
void bruteGPU(std::vector<std::vector<double>> &columns, double min, double max, int nbins, std::vector<double> &xi,
              long chunksize, BruteArithmetic arithmetic = BruteFP32);

}  // namespace turbooctospice

#endif
