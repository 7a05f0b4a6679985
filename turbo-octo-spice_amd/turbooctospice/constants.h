// constants.h -- physical and angular constants (same values as the
// reference's turbooctospice/constants.h:10-14).
#ifndef TOS_AMD_TURBOOCTOSPICE_CONSTANTS
#define TOS_AMD_TURBOOCTOSPICE_CONSTANTS

#include <cmath>

namespace turbooctospice {
const double lyA = 1216.0;
const double logLyA = std::log10(lyA);
const double pi = std::atan(1.0) * 4;
const double deg2rad = pi / 180.0;
const double rad2arcmin = 60.0 / deg2rad;
}  // namespace turbooctospice

#endif
