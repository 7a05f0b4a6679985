"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  See oracle.h for the
reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

POLAR, CARTESIAN, OBSERVING = 0, 1, 2
SKY_HEALPIX, SKY_PLATES = 0, 1


class XiBin(C.Structure):
    _fields_ = [("didj", C.c_double), ("di", C.c_double), ("dj", C.c_double),
                ("wgt", C.c_double), ("wi", C.c_double), ("wj", C.c_double),
                ("z", C.c_double), ("num_pairs", C.c_int64)]


XIBIN_DTYPE = np.dtype([("didj", "f8"), ("di", "f8"), ("dj", "f8"), ("wgt", "f8"),
                        ("wi", "f8"), ("wj", "f8"), ("z", "f8"), ("num_pairs", "i8")])


class Sightlines(C.Structure):
    _fields_ = [("n_los", C.c_int64), ("offsets", C.POINTER(C.c_int64)),
                ("ra", C.POINTER(C.c_double)), ("dec", C.POINTER(C.c_double)),
                ("value", C.POINTER(C.c_float)), ("loglam", C.POINTER(C.c_float)),
                ("weight", C.POINTER(C.c_float)), ("distance", C.POINTER(C.c_float))]


class Grid(C.Structure):
    _fields_ = [("type", C.c_int), ("nbins", C.c_int32 * 3), ("min", C.c_double * 3),
                ("width", C.c_double * 3), ("max", C.c_double * 3)]


class SkyBins(C.Structure):
    _fields_ = [("kind", C.c_int), ("bin_of_los", C.POINTER(C.c_int32)),
                ("n_plates", C.c_int64), ("plate_id", C.POINTER(C.c_int32)),
                ("plate_ra", C.POINTER(C.c_double)), ("plate_dec", C.POINTER(C.c_double))]


class Stats(C.Structure):
    _fields_ = [("num_sightline_pairs_used", C.c_uint64), ("num_pixel_pairs", C.c_uint64),
                ("num_pixel_pairs_used", C.c_uint64), ("num_pixels", C.c_uint64),
                ("num_sightline_pairs_considered", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_cvt_u32.restype = C.c_uint32
        L.ora_cvt_u32.argtypes = [C.c_double]
        L.ora_cos_separation.restype = C.c_double
        L.ora_cos_separation.argtypes = [C.c_double] * 4
        L.ora_max_angular_scale.restype = C.c_double
        L.ora_max_angular_scale.argtypes = [C.POINTER(Grid), C.c_double]
        L.ora_accumulate_pair.restype = None
        L.ora_accumulate_pair.argtypes = [C.POINTER(XiBin)] + [C.c_float] * 6
        L.ora_finalize.restype = None
        L.ora_finalize.argtypes = [C.POINTER(XiBin)]
        L.ora_accumulate_los_pair.restype = C.c_uint64
        L.ora_accumulate_los_pair.argtypes = [C.POINTER(Sightlines), C.POINTER(Grid), C.c_int64,
                                              C.c_int64, C.c_double, C.c_void_p]
        L.ora_xi_run.restype = C.c_int64
        L.ora_xi_run.argtypes = [C.POINTER(Sightlines), C.POINTER(Grid), C.c_double,
                                 C.POINTER(SkyBins), C.c_int, C.c_int64, C.c_int64,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(Stats)]
        L.ora_ang2pix_ring.restype = C.c_int64
        L.ora_ang2pix_ring.argtypes = [C.c_int64, C.c_double, C.c_double]
        L.ora_pix2ang_ring.restype = None
        L.ora_pix2ang_ring.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.ora_radec2pix.restype = C.c_int64
        L.ora_radec2pix.argtypes = [C.c_int64, C.c_double, C.c_double]
        L.ora_brute3d_f64.restype = None
        L.ora_brute3d_f64.argtypes = [C.c_void_p] * 5 + [C.c_int64, C.c_double, C.c_double,
                                                         C.c_int32] + [C.c_void_p] * 3
        L.ora_brute3d_f32.restype = None
        L.ora_brute3d_f32.argtypes = [C.c_void_p] * 5 + [C.c_int64, C.c_float, C.c_float,
                                                         C.c_int32] + [C.c_void_p] * 3
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_grid(gspec) -> Grid:
    """gspec: object with .type, .nbins, .min, .width, .max (3-sequences)."""
    g = Grid()
    g.type = int(gspec.type)
    for k in range(3):
        g.nbins[k] = int(gspec.nbins[k])
        g.min[k] = float(gspec.min[k])
        g.width[k] = float(gspec.width[k])
        g.max[k] = float(gspec.max[k])
    return g


class _SightlineArrays:
    """Holds contiguous copies so the ctypes struct's pointers stay valid."""

    def __init__(self, cat):
        self.offsets = np.ascontiguousarray(cat.offsets, dtype=np.int64)
        self.ra = np.ascontiguousarray(cat.ra, dtype=np.float64)
        self.dec = np.ascontiguousarray(cat.dec, dtype=np.float64)
        self.value = np.ascontiguousarray(cat.value, dtype=np.float32)
        self.loglam = np.ascontiguousarray(cat.loglam, dtype=np.float32)
        self.weight = np.ascontiguousarray(cat.weight, dtype=np.float32)
        self.distance = np.ascontiguousarray(cat.distance, dtype=np.float32)
        s = Sightlines()
        s.n_los = len(self.ra)
        s.offsets = _ptr(self.offsets, C.c_int64)
        s.ra = _ptr(self.ra, C.c_double)
        s.dec = _ptr(self.dec, C.c_double)
        s.value = _ptr(self.value, C.c_float)
        s.loglam = _ptr(self.loglam, C.c_float)
        s.weight = _ptr(self.weight, C.c_float)
        s.distance = _ptr(self.distance, C.c_float)
        self.s = s


def healpix_bins(cat, order: int) -> np.ndarray:
    nside = 1 << order
    L = lib()
    return np.array([L.ora_radec2pix(nside, float(r), float(d)) for r, d in zip(cat.ra, cat.dec)],
                    dtype=np.int32)


def xi_run(cat, gspec, scale: float, *, order: int | None = 4, plates=None, nthreads: int = 8,
           task_stride: int = 1, task_phase: int = 0, per_skybin: bool = False,
           bin_of_los=None):
    """XiEstimator::run restated on the CPU.

    Returns dict(final=XiBin array, raw=XiBin array, stats=dict,
                 skybin_keys=..., per_skybin=[nsky, nbins] XiBin array or None).
    plates: None for HEALPix binning, else (bin_of_los, plate_id, plate_ra, plate_dec).
    """
    L = lib()
    arrs = _SightlineArrays(cat)
    g = make_grid(gspec)
    nbins = int(np.prod(gspec.nbins))
    sky = SkyBins()
    keep = []
    if plates is None:
        if bin_of_los is None:
            bin_of_los = healpix_bins(cat, order)
        bol = np.ascontiguousarray(bin_of_los, dtype=np.int32)
        keep.append(bol)
        sky.kind = SKY_HEALPIX
        sky.bin_of_los = _ptr(bol, C.c_int32)
        sky.n_plates = 0
    else:
        bol, pid, pra, pdec = plates
        bol = np.ascontiguousarray(bol, dtype=np.int32)
        pid = np.ascontiguousarray(pid, dtype=np.int32)
        pra = np.ascontiguousarray(pra, dtype=np.float64)
        pdec = np.ascontiguousarray(pdec, dtype=np.float64)
        keep += [bol, pid, pra, pdec]
        sky.kind = SKY_PLATES
        sky.bin_of_los = _ptr(bol, C.c_int32)
        sky.n_plates = len(pid)
        sky.plate_id = _ptr(pid, C.c_int32)
        sky.plate_ra = _ptr(pra, C.c_double)
        sky.plate_dec = _ptr(pdec, C.c_double)
    n_keys = len(np.unique(bol))
    final = np.zeros(nbins, dtype=XIBIN_DTYPE)
    raw = np.zeros(nbins, dtype=XIBIN_DTYPE)
    keys = np.zeros(max(n_keys, 1), dtype=np.int32)
    per = np.zeros((n_keys, nbins), dtype=XIBIN_DTYPE) if per_skybin else None
    st = Stats()
    n = L.ora_xi_run(C.byref(arrs.s), C.byref(g), float(scale), C.byref(sky), int(nthreads),
                     int(task_stride), int(task_phase), final.ctypes.data, raw.ctypes.data,
                     per.ctypes.data if per is not None else None, keys.ctypes.data,
                     C.byref(st))
    if n < 0:
        raise RuntimeError("ora_xi_run rejected its input")
    return {"final": final, "raw": raw, "stats": st.as_dict(), "n_tasks": int(n),
            "skybin_keys": keys[:n], "per_skybin": per[:n] if per is not None else None}


def accumulate_los_pair(cat, gspec, primary: int, other: int, cos_sep: float):
    L = lib()
    arrs = _SightlineArrays(cat)
    g = make_grid(gspec)
    xi = np.zeros(int(np.prod(gspec.nbins)), dtype=XIBIN_DTYPE)
    used = L.ora_accumulate_los_pair(C.byref(arrs.s), C.byref(g), int(primary), int(other),
                                     float(cos_sep), xi.ctypes.data)
    return xi, int(used)


def brute3d_f64(x, y, z, d, w, mn, mx, nbins):
    L = lib()
    cols = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, d, w)]
    ds = np.zeros(nbins); ws = np.zeros(nbins); cn = np.zeros(nbins, dtype=np.int64)
    L.ora_brute3d_f64(*[c.ctypes.data for c in cols], len(cols[0]), float(mn), float(mx),
                      int(nbins), ds.ctypes.data, ws.ctypes.data, cn.ctypes.data)
    return ds, ws, cn


def brute3d_f32(x, y, z, d, w, mn, mx, nbins):
    L = lib()
    cols = [np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, z, d, w)]
    ds = np.zeros(nbins + 2); ws = np.zeros(nbins + 2); cn = np.zeros(nbins + 2, dtype=np.int64)
    L.ora_brute3d_f32(*[c.ctypes.data for c in cols], len(cols[0]), float(mn), float(mx),
                      int(nbins), ds.ctypes.data, ws.ctypes.data, cn.ctypes.data)
    return ds, ws, cn


def cvt_u32(x: float) -> int:
    return int(lib().ora_cvt_u32(float(x)))


def cos_separation(ra1, dec1, ra2, dec2) -> float:
    return float(lib().ora_cos_separation(ra1, dec1, ra2, dec2))


def max_angular_scale(gspec, scale) -> float:
    g = make_grid(gspec)
    return float(lib().ora_max_angular_scale(C.byref(g), float(scale)))


def ang2pix_ring(nside, theta, phi) -> int:
    return int(lib().ora_ang2pix_ring(int(nside), float(theta), float(phi)))


def pix2ang_ring(nside, pix):
    t = C.c_double(); p = C.c_double()
    lib().ora_pix2ang_ring(int(nside), int(pix), C.byref(t), C.byref(p))
    return t.value, p.value
