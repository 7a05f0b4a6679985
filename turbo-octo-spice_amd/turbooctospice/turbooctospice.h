// turbooctospice.h -- umbrella header of the C++ API (the reference's
// turbooctospice/turbooctospice.h).
#ifndef TOS_AMD_TURBOOCTOSPICE
#define TOS_AMD_TURBOOCTOSPICE

#include "AbsTwoPointGrid.h"
#include "Cosmology.h"
#include "HDF5Delta.h"
#include "RuntimeError.h"
#include "SkyBins.h"
#include "ThreadPool.h"
#include "XiEstimator.h"
#include "constants.h"
#include "types.h"

#endif
