"""ctypes binding of libtos_amd.so (include/tos_amd.h).

This is the Python-side mirror of the C-ABI, used by the tests, bench.py and
__graft_entry__.  It loads only the in-tree library and raises if it is
missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# TOS_AMD_LIB: path of another build of the library (default: the in-tree one)
LIB_PATH = os.environ.get("TOS_AMD_LIB") or os.path.join(PKG_DIR, "_build", "libtos_amd.so")

TOS_OK = 0
STATUS_NAMES = {0: "TOS_OK", 1: "TOS_ERR_INVALID", 2: "TOS_ERR_NOMEM", 3: "TOS_ERR_DEVICE",
                4: "TOS_ERR_INTERNAL", 5: "TOS_ERR_NOT_READY", 6: "TOS_ERR_CANCELLED",
                7: "TOS_ERR_TIMEOUT"}

POLAR, CARTESIAN, OBSERVING = 0, 1, 2
SKY_HEALPIX, SKY_PLATES = 0, 1
FP64, FP32C = 0, 1
ACC_CORE, ACC_ALL = 0, 1
BRUTE_FP64, BRUTE_FP32 = 0, 1
XI_NO_WINDOW, XI_NO_TRANSPOSE, XI_NO_INTERLEAVE, XI_NO_ZCONST = 1, 2, 4, 8


class TosError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Grid(C.Structure):
    _fields_ = [("type", C.c_int32), ("nbins", C.c_int32 * 3), ("min", C.c_double * 3),
                ("width", C.c_double * 3), ("max", C.c_double * 3)]


class Sightlines(C.Structure):
    _fields_ = [("n_los", C.c_int64), ("offsets", C.c_void_p), ("ra", C.c_void_p),
                ("dec", C.c_void_p), ("value", C.c_void_p), ("loglam", C.c_void_p),
                ("weight", C.c_void_p), ("distance", C.c_void_p)]


class SkyBinsDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("order", C.c_int32), ("plate_of_los", C.c_void_p),
                ("n_plates", C.c_int64), ("plate_id", C.c_void_p), ("plate_ra", C.c_void_p),
                ("plate_dec", C.c_void_p)]


class PairStats(C.Structure):
    _fields_ = [("num_sightline_pairs", C.c_uint64), ("num_sightline_pairs_used", C.c_uint64),
                ("num_pixel_pairs", C.c_uint64), ("num_pixels", C.c_uint64),
                ("n_rows", C.c_int64), ("n_skybins", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class PairCSR(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("row_los", C.c_void_p), ("row_skybin", C.c_void_p),
                ("row_ptr", C.c_void_p), ("other", C.c_void_p), ("cos_sep", C.c_void_p)]


class XiOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("accum", C.c_int32), ("per_skybin", C.c_int32),
                ("block_size", C.c_int32), ("blocks_per_cu", C.c_int32), ("flags", C.c_int32),
                ("wg_per_slot", C.c_int32), ("reserved", C.c_int32)]


class XiResult(C.Structure):
    _fields_ = [("num_pairs", C.c_void_p), ("didj", C.c_void_p), ("wgt", C.c_void_p),
                ("di", C.c_void_p), ("dj", C.c_void_p), ("wi", C.c_void_p), ("wj", C.c_void_p),
                ("z", C.c_void_p), ("skybin_keys", C.c_void_p), ("sub_num_pairs", C.c_void_p),
                ("sub_didj", C.c_void_p), ("sub_wgt", C.c_void_p),
                ("num_pixel_pairs_used", C.c_uint64), ("num_guard_rechecks", C.c_uint64),
                ("num_pairs_evaluated", C.c_uint64), ("kernel_ms", C.c_double)]


class BatchJob(C.Structure):
    _fields_ = [("cat", C.c_void_p), ("pairs", C.c_void_p), ("grid", Grid),
                ("precision", C.c_int32), ("accum", C.c_int32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("wavefront_size", C.c_int32), ("lds_per_block", C.c_int64),
                ("global_mem", C.c_int64), ("clock_khz", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_inflight", C.c_int32),
                ("nthreads", C.c_int32), ("search_on_device", C.c_int32), ("reserved", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("sightlines", Sightlines), ("sky", SkyBinsDesc), ("grid", Grid),
                ("scale", C.c_double), ("precision", C.c_int32)]


class JobTiming(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("admit_ms", C.c_double), ("start_ms", C.c_double),
                ("done_ms", C.c_double), ("footprint_bytes", C.c_int64), ("status", C.c_int32),
                ("num_pixel_pairs", C.c_uint64)]


# every symbol declared in include/tos_amd.h, with (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "tos_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "tos_ctx_destroy": (C.c_int, [_P]),
    "tos_last_error": (C.c_char_p, [_P]),
    "tos_device_info_get": (C.c_int, [_P, C.POINTER(DeviceInfo)]),
    "tos_build_info": (C.c_char_p, []),
    "tos_grid_parse_axis": (C.c_int, [C.c_char_p, C.POINTER(Grid), C.c_int]),
    "tos_grid_max_angular_scale": (C.c_double, [C.POINTER(Grid), C.c_double]),
    "tos_grid_bin_centers": (C.c_int, [C.POINTER(Grid), C.c_int32, C.POINTER(C.c_double)]),
    "tos_catalog_upload": (C.c_int, [_P, C.POINTER(Sightlines), C.POINTER(_P)]),
    "tos_catalog_free": (C.c_int, [_P, _P]),
    "tos_catalog_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]),
    "tos_healpix_radec2pix": (C.c_int, [C.c_int32, C.c_int64, _P, _P, _P]),
    "tos_pairs_build": (C.c_int, [_P, _P, C.POINTER(SkyBinsDesc), C.c_double, C.c_int, C.c_int,
                                  C.POINTER(_P)]),
    "tos_pairs_upload": (C.c_int, [_P, _P, C.POINTER(PairCSR), C.POINTER(_P)]),
    "tos_pairs_free": (C.c_int, [_P, _P]),
    "tos_pairs_stats": (C.c_int, [_P, C.POINTER(PairStats)]),
    "tos_pairs_download": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "tos_xi_run": (C.c_int, [_P, _P, _P, C.POINTER(Grid), C.POINTER(XiOpts),
                             C.POINTER(XiResult)]),
    "tos_batch_run": (C.c_int, [_P, C.c_int32, _P, C.c_int, _P, C.POINTER(C.c_double)]),
    "tos_xi_autotune": (C.c_int, [_P, _P, _P, C.POINTER(Grid), C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "tos_xi_tuned_shape": (C.c_int, [_P, C.POINTER(Grid), C.c_int32, C.POINTER(C.c_int32 * 4)]),
    "tos_server_start": (C.c_int, [_P, C.POINTER(ServerOpts), C.POINTER(_P)]),
    "tos_server_stop": (C.c_int, [_P]),
    "tos_job_submit": (C.c_int, [_P, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    "tos_job_poll": (C.c_int, [_P, C.c_uint64]),
    "tos_job_wait": (C.c_int, [_P, C.c_uint64, C.c_double]),
    "tos_job_cancel": (C.c_int, [_P, C.c_uint64]),
    "tos_job_result": (C.c_int, [_P, C.c_uint64, C.POINTER(XiResult), C.POINTER(JobTiming)]),
    "tos_job_timing_get": (C.c_int, [_P, C.c_uint64, C.POINTER(JobTiming)]),
    "tos_xi_run_noalloc": (C.c_int, [_P, _P, _P, C.POINTER(Grid), C.c_int32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "tos_job_footprint": (C.c_int, [C.POINTER(JobDesc), C.POINTER(C.c_int64)]),
    "tos_brute3d": (C.c_int, [_P, C.c_int32, C.c_int64, _P, _P, _P, _P, _P, C.c_double,
                              C.c_double, C.c_int32, _P, _P, _P, C.POINTER(C.c_double)]),
    "tos_covariance_check": (C.c_int, [C.c_int32, _P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "tos_subsample_covariance": (C.c_int, [_P, C.c_int64, C.c_int32, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    "tos_cosmo_distances": (C.c_int, [C.c_double, C.c_double, C.c_int64, _P, _P, _P]),
    "tos_pixel_distances": (C.c_int, [_P, C.c_double, C.c_double, C.c_int64, _P, _P]),
    "tos_selftest_f64": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P, _P]),
}

_lib = None


def lib():
    """Load the in-tree library (built by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libtos_amd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None, what=""):
    if status != TOS_OK:
        msg = ""
        if ctx is not None:
            m = lib().tos_last_error(ctx)
            msg = m.decode() if m else ""
        raise TosError(status, f"{what}: {msg}" if what else msg)
