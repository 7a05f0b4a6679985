// HDF5Delta.h -- reads delta-field sightlines from an HDF5 file (mirrors
// the reference's HDF5Delta.h:12-24 and the file layout of HDF5Delta.cc).
#ifndef TOS_AMD_TURBOOCTOSPICE_HDF5_DELTA
#define TOS_AMD_TURBOOCTOSPICE_HDF5_DELTA

#include <string>
#include <vector>

#include "types.h"

namespace turbooctospice {

class HDF5Delta {
public:
    explicit HDF5Delta(std::string filename);
    /// Reads every subgroup of `groupname` in name order: attributes ra,
    /// dec, z (double) and plate (int); datasets delta, weight, r_comov and
    /// loglam (float).  Pixels with non-finite delta or weight, or weight <= 0,
    /// are dropped; sightlines left without pixels are skipped.  forest_id
    /// counts loaded sightlines across calls (HDF5Delta.cc:51).
    std::vector<Forest> loadForests(std::string groupname);
    /// Writes sightlines in the same layout (group per sightline, named by
    /// a zero-padded index so that name order is load order).
    static void saveForests(std::string const &filename, std::string const &groupname,
                            std::vector<Forest> const &forests, std::vector<double> const &redshifts);
private:
    std::string _filename;
};

}  // namespace turbooctospice

#endif
