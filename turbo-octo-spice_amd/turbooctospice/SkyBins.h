// SkyBins.h -- sky partitioning used to find candidate sightline pairs
// (mirrors the reference's SkyBins.h:16-63, HealpixBins.h:18-82 and
// PlateBins.h:18-92: same interface, items keyed by sky bin, occupied bins
// in first-insertion order, bin contents in insertion order).
#ifndef TOS_AMD_TURBOOCTOSPICE_SKY_BINS
#define TOS_AMD_TURBOOCTOSPICE_SKY_BINS

#include <map>
#include <string>
#include <vector>

#include "constants.h"
#include "types.h"

namespace turbooctospice {

template <typename T> class SkyBins {
public:
    typedef std::map<int, std::vector<T>> Bins;
    SkyBins() : num_entries_(0) {}
    virtual ~SkyBins() {}
    /// Adds an item to the bin of the forest's sky position.
    virtual void addItem(const Forest &forest, const T &item) = 0;
    std::vector<int> getOccupiedBins() const { return occupied_bins_; }
    /// Bins that may hold items within `radius` (radians) of (ra, dec): a
    /// superset of the bins holding such items.
    virtual std::vector<int> getBinIndicesWithinRadius(double ra, double dec, double radius) const = 0;
    bool checkBinExists(int index) const { return bins_.find(index) != bins_.end(); }
    const typename Bins::mapped_type &getBinContents(const typename Bins::key_type &k) const { return bins_.at(k); }
    int getNBins() const { return (int)bins_.size(); }
    unsigned long getNEntries() const { return num_entries_; }
    virtual void printBinCenter(const int index) const = 0;
    /// Sky bin of a forest (its HEALPix cell or plate).
    virtual int binOf(const Forest &forest) const = 0;
protected:
    void insert(int key, const T &item)
    {
        auto it = bins_.find(key);
        if (it == bins_.end()) {
            bins_[key] = std::vector<T>(1, item);
            occupied_bins_.push_back(key);
        } else {
            it->second.push_back(item);
        }
        ++num_entries_;
    }
    Bins bins_;
    unsigned long num_entries_;
    std::vector<int> occupied_bins_;
};

typedef SkyBins<int> SkyBinsI;

/// HEALPix RING cells of nside = 2^order (HealpixBins.h:18-82).
class HealpixBinsI : public SkyBins<int> {
public:
    explicit HealpixBinsI(int order);
    void addItem(const Forest &forest, const int &item) override;
    std::vector<int> getBinIndicesWithinRadius(double ra, double dec, double radius) const override;
    int ang2pix(double ra, double dec) const;
    void printBinCenter(const int index) const override;
    int binOf(const Forest &forest) const override { return ang2pix(forest.ra, forest.dec); }
    int order() const { return order_; }
private:
    int order_;
    long nside_;
    // per occupied cell: the largest angle between the cell centre and a
    // member, so that a query returns every cell that can hold a member
    // within the radius (triangle inequality)
    std::map<int, double> cell_radius_;
};

/// Sightlines binned by their plate (PlateBins.h:18-92).  The neighbour rule
/// is the reference's: plates whose centre lies within `radius` of the
/// 1.5-degree annulus around the query point.
class PlateBinsI : public SkyBins<int> {
public:
    typedef std::map<int, SkyObject> PlateList;
    /// Plate list from a FITS table with PLATE, RACEN, DECCEN columns (degrees),
    /// as the reference's platelist.fits (PlateBins.h:23-50), or from a text
    /// file with one "plate ra_deg dec_deg" line per plate.
    explicit PlateBinsI(std::string const &platelist_filename);
    explicit PlateBinsI(PlateList const &plates);
    void addItem(const Forest &forest, const int &item) override;
    std::vector<int> getBinIndicesWithinRadius(double ra, double dec, double radius) const override;
    void printBinCenter(const int index) const override;
    int binOf(const Forest &forest) const override { return forest.plate; }
    PlateList const &plates() const { return platelist_; }
private:
    PlateList platelist_;
};

}  // namespace turbooctospice

#endif
