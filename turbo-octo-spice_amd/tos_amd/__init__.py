"""tos_amd -- Python front end of the gfx950 pair-counting correlation
estimator (libtos_amd.so).  The compute path is the HIP library only."""
from . import _lib
from ._lib import (ACC_ALL, ACC_CORE, BRUTE_FP32, BRUTE_FP64, CARTESIAN, FP32C, FP64, OBSERVING,
                   POLAR, XI_NO_INTERLEAVE, XI_NO_TRANSPOSE, XI_NO_WINDOW, XI_NO_ZCONST, TosError, lib)
from .grids import GridSpec, parse_binning
from .estimator import Device, DeviceCatalog, DevicePairs, XiHistogram, covariance_check, healpix_radec2pix
from . import synth

__all__ = ["GridSpec", "parse_binning", "Device", "DeviceCatalog", "DevicePairs", "XiHistogram",
           "covariance_check", "healpix_radec2pix", "synth", "lib", "TosError", "POLAR", "CARTESIAN", "OBSERVING",
           "FP64", "FP32C", "ACC_CORE", "ACC_ALL", "BRUTE_FP64", "BRUTE_FP32"]
