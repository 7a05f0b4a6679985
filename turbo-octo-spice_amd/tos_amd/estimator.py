"""Python front end of the C-ABI: device context, resident catalogues,
candidate pairs and the estimator run (the same operations as the C++
XiEstimator in turbooctospice/, for tests and benchmarks)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib


@dataclass
class XiHistogram:
    num_pairs: np.ndarray     # uint64 per bin
    didj: np.ndarray          # raw sum of w1 d1 w2 d2
    wgt: np.ndarray           # raw sum of w1 w2
    extra: dict               # di, dj, wi, wj, z with ACC_ALL
    num_pixel_pairs_used: int
    num_guard_rechecks: int
    kernel_ms: float
    num_pairs_evaluated: int = 0

    def xi(self) -> np.ndarray:
        """XiBin::finalize for didj (types.h:97)."""
        out = np.zeros_like(self.didj)
        m = self.wgt > 0
        out[m] = self.didj[m] / self.wgt[m]
        return out


class Device:
    """One tos_ctx on a HIP device."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().tos_ctx_create(int(device), C.byref(h))
        if st != 0:
            raise _lib.TosError(st, f"tos_ctx_create({device}) failed")
        self.h = h

    def close(self):
        if self.h:
            lib().tos_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        d = _lib.DeviceInfo()
        check(lib().tos_device_info_get(self.h, C.byref(d)), self.h, "tos_device_info_get")
        return {"name": d.name.decode(), "arch": d.arch.decode(), "compute_units": d.compute_units,
                "wavefront_size": d.wavefront_size, "lds_per_block": d.lds_per_block,
                "global_mem": d.global_mem, "clock_khz": d.clock_khz}

    # -- catalogue and pairs -------------------------------------------------
    def upload(self, cat) -> "DeviceCatalog":
        s, keep = cat.to_c()
        h = C.c_void_p()
        check(lib().tos_catalog_upload(self.h, C.byref(s), C.byref(h)), self.h, "tos_catalog_upload")
        return DeviceCatalog(self, h, cat)

    def brute3d(self, x, y, z, d, w, mn, mx, nbins, mode=_lib.BRUTE_FP64):
        cols = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, d, w)]
        nslot = nbins if mode == _lib.BRUTE_FP64 else nbins + 2
        ds = np.zeros(nslot); ws = np.zeros(nslot); cn = np.zeros(nslot, dtype=np.uint64)
        ms = C.c_double()
        check(lib().tos_brute3d(self.h, int(mode), len(cols[0]), *[c.ctypes.data for c in cols],
                                float(mn), float(mx), int(nbins), ds.ctypes.data, ws.ctypes.data,
                                cn.ctypes.data, C.byref(ms)), self.h, "tos_brute3d")
        return ds, ws, cn, ms.value

    def pixel_distances(self, loglam, omega_lambda=0.7, omega_matter=0.3):
        """The per-pixel distance transform on the device
        (tos_pixel_distances): float32 comoving distances, Mpc/h."""
        loglam = np.ascontiguousarray(loglam, dtype=np.float32)
        out = np.empty_like(loglam)
        check(lib().tos_pixel_distances(self.h, float(omega_lambda), float(omega_matter), len(loglam),
                                        loglam.ctypes.data, out.ctypes.data), self.h, "tos_pixel_distances")
        return out

    def subsample_covariance(self, didj, wgt):
        """Weighted covariance of sky-bin subsamples (tos_subsample_covariance,
        XiEstimator.cc:60-97): didj, wgt are [n_samples, nbins].  Returns
        (cov [nbins, nbins], mean [nbins], total weight)."""
        didj = np.ascontiguousarray(didj, dtype=np.float64)
        wgt = np.ascontiguousarray(wgt, dtype=np.float64)
        if didj.shape != wgt.shape or didj.ndim != 2:
            raise ValueError("didj and wgt must both be [n_samples, nbins]")
        k, nb = didj.shape
        cov = np.zeros((nb, nb))
        mean = np.zeros(nb)
        ws = C.c_double()
        check(lib().tos_subsample_covariance(self.h, k, nb, didj.ctypes.data, wgt.ctypes.data, cov.ctypes.data,
                                             mean.ctypes.data, C.byref(ws)), self.h, "tos_subsample_covariance")
        return cov, mean, ws.value

    def batch_run(self, jobs, use_graph=False):
        """Run many (pairs, grid, precision) jobs in one batched launch
        (tos_batch_run).  Returns ([XiHistogram per job], batch_ms)."""
        n = len(jobs)
        arr = (_lib.BatchJob * max(1, n))()
        outs, keep = [], []
        res = (_lib.XiResult * max(1, n))()
        for k, (pairs, grid, prec) in enumerate(jobs):
            arr[k].cat = pairs.dcat.h
            arr[k].pairs = pairs.h
            arr[k].grid = grid.to_c()
            arr[k].precision = int(prec)
            arr[k].accum = _lib.ACC_CORE
            nb = grid.nbins_total
            cnt, didj, wgt = np.zeros(nb, np.uint64), np.zeros(nb), np.zeros(nb)
            keep.append((cnt, didj, wgt))
            res[k].num_pairs = cnt.ctypes.data
            res[k].didj = didj.ctypes.data
            res[k].wgt = wgt.ctypes.data
        ms = C.c_double(0)
        check(lib().tos_batch_run(self.h, n, C.cast(arr, C.c_void_p), int(bool(use_graph)),
                                  C.cast(res, C.c_void_p), C.byref(ms)), self.h, "tos_batch_run")
        for k in range(n):
            cnt, didj, wgt = keep[k]
            outs.append(XiHistogram(cnt, didj, wgt, {}, int(res[k].num_pixel_pairs_used),
                                    int(res[k].num_guard_rechecks), float(res[k].kernel_ms),
                                    int(res[k].num_pairs_evaluated)))
        return outs, float(ms.value)

    def autotune(self, pairs, grid, precision=_lib.FP32C):
        """Pick the fastest launch shape for this kernel variant
        (tos_xi_autotune); returns (block, blocks_per_cu, kernel_ms)."""
        b, bpc, ms = C.c_int32(0), C.c_int32(0), C.c_double(0)
        g = grid.to_c()
        check(lib().tos_xi_autotune(self.h, pairs.dcat.h, pairs.h, C.byref(g), int(precision), C.byref(b),
                                    C.byref(bpc), C.byref(ms)), self.h, "tos_xi_autotune")
        return int(b.value), int(bpc.value), float(ms.value)

    def tuned_shape(self, grid, precision=_lib.FP32C):
        """The kept shape (tos_xi_tuned_shape): dict of block,
        blocks_per_cu, wg_per_slot, step_cap (0: the default)."""
        s = (C.c_int32 * 4)()
        g = grid.to_c()
        check(lib().tos_xi_tuned_shape(self.h, C.byref(g), int(precision), C.byref(s)), self.h,
              "tos_xi_tuned_shape")
        return dict(zip(("block", "blocks_per_cu", "wg_per_slot", "step_cap"), (int(v) for v in s)))

    def selftest_f64(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        s = np.empty_like(a); q = np.empty_like(a); c = np.empty(len(a), dtype=np.uint32)
        check(lib().tos_selftest_f64(self.h, len(a), a.ctypes.data, b.ctypes.data, s.ctypes.data,
                                     q.ctypes.data, c.ctypes.data), self.h, "tos_selftest_f64")
        return s, q, c


def sky_desc(order=4, plates=None):
    """tos_skybins for HEALPix (order) or plates = (plate_of_los, ids, ra, dec)."""
    d = _lib.SkyBinsDesc()
    keep = []
    if plates is None:
        d.kind = _lib.SKY_HEALPIX
        d.order = int(order)
    else:
        pol, pid, pra, pdec = [np.ascontiguousarray(a, dtype=t) for a, t in
                               zip(plates, (np.int32, np.int32, np.float64, np.float64))]
        keep += [pol, pid, pra, pdec]
        d.kind = _lib.SKY_PLATES
        d.plate_of_los = pol.ctypes.data
        d.n_plates = len(pid)
        d.plate_id = pid.ctypes.data
        d.plate_ra = pra.ctypes.data
        d.plate_dec = pdec.ctypes.data
    return d, keep


class DeviceCatalog:
    def __init__(self, dev: Device, h, cat):
        self.dev, self.h, self.cat = dev, h, cat

    def free(self):
        if self.h and self.dev.h:
            lib().tos_catalog_free(self.dev.h, self.h)
        self.h = None

    def pairs(self, max_ang: float, order: int = 4, plates=None, nthreads: int = 8,
              on_device: bool = False) -> "DevicePairs":
        d, keep = sky_desc(order, plates)
        h = C.c_void_p()
        check(lib().tos_pairs_build(self.dev.h, self.h, C.byref(d), float(max_ang), int(nthreads),
                                    int(bool(on_device)), C.byref(h)), self.dev.h, "tos_pairs_build")
        return DevicePairs(self, h)

    def pairs_from_csr(self, row_los, row_ptr, other, cos_sep, row_skybin=None) -> "DevicePairs":
        arrs = [np.ascontiguousarray(row_los, dtype=np.int32), np.ascontiguousarray(row_ptr, dtype=np.int64),
                np.ascontiguousarray(other, dtype=np.int32), np.ascontiguousarray(cos_sep, dtype=np.float64)]
        rs = None if row_skybin is None else np.ascontiguousarray(row_skybin, dtype=np.int32)
        csr = _lib.PairCSR()
        csr.n_rows = len(arrs[0])
        csr.row_los, csr.row_ptr, csr.other, csr.cos_sep = [a.ctypes.data for a in arrs]
        csr.row_skybin = rs.ctypes.data if rs is not None else None
        h = C.c_void_p()
        check(lib().tos_pairs_upload(self.dev.h, self.h, C.byref(csr), C.byref(h)), self.dev.h,
              "tos_pairs_upload")
        return DevicePairs(self, h)


class DevicePairs:
    def __init__(self, dcat: DeviceCatalog, h):
        self.dcat, self.h = dcat, h

    def free(self):
        if self.h and self.dcat.dev.h:
            lib().tos_pairs_free(self.dcat.dev.h, self.h)
        self.h = None

    def stats(self) -> dict:
        s = _lib.PairStats()
        check(lib().tos_pairs_stats(self.h, C.byref(s)))
        return s.as_dict()

    def download(self):
        st = self.stats()
        nr = st["n_rows"]
        npairs = st["num_sightline_pairs_used"]
        row_los = np.zeros(nr, np.int32); row_sky = np.zeros(nr, np.int32)
        row_ptr = np.zeros(nr + 1, np.int64); other = np.zeros(npairs, np.int32)
        cs = np.zeros(npairs, np.float64)
        check(lib().tos_pairs_download(self.dcat.dev.h, self.h, row_los.ctypes.data, row_sky.ctypes.data,
                                       row_ptr.ctypes.data, other.ctypes.data, cs.ctypes.data),
              self.dcat.dev.h, "tos_pairs_download")
        return row_los, row_sky, row_ptr, other, cs

    def run_per_skybin(self, grid, precision=_lib.FP64, flags=0):
        """One pass with a histogram per occupied sky bin (XiEstimator.h:78).
        Returns (total XiHistogram, keys, sub_num_pairs, sub_didj, sub_wgt),
        the sub-arrays [n_skybins, nbins] in ascending key order."""
        nb = grid.nbins_total
        nk = self.stats()["n_skybins"]
        cnt, didj, wgt = np.zeros(nb, np.uint64), np.zeros(nb), np.zeros(nb)
        keys = np.zeros(max(nk, 1), np.int32)
        sn = np.zeros((max(nk, 1), nb), np.uint64)
        sd = np.zeros((max(nk, 1), nb))
        sw = np.zeros((max(nk, 1), nb))
        res = _lib.XiResult()
        res.num_pairs, res.didj, res.wgt = cnt.ctypes.data, didj.ctypes.data, wgt.ctypes.data
        res.skybin_keys = keys.ctypes.data
        res.sub_num_pairs, res.sub_didj, res.sub_wgt = sn.ctypes.data, sd.ctypes.data, sw.ctypes.data
        opts = _lib.XiOpts()
        opts.precision = int(precision)
        opts.accum = _lib.ACC_CORE
        opts.per_skybin = 1
        opts.flags = int(flags)
        g = grid.to_c()
        dev = self.dcat.dev
        check(lib().tos_xi_run(dev.h, self.dcat.h, self.h, C.byref(g), C.byref(opts), C.byref(res)),
              dev.h, "tos_xi_run(per_skybin)")
        h = XiHistogram(cnt, didj, wgt, {}, int(res.num_pixel_pairs_used), int(res.num_guard_rechecks),
                        float(res.kernel_ms), int(res.num_pairs_evaluated))
        return h, keys[:nk], sn[:nk], sd[:nk], sw[:nk]

    def run(self, grid, precision=_lib.FP64, accum=_lib.ACC_CORE, block_size=0,
            blocks_per_cu=0, flags=0, wg_per_slot=0) -> XiHistogram:
        nb = grid.nbins_total
        cnt = np.zeros(nb, np.uint64)
        sums = {k: np.zeros(nb) for k in ("didj", "wgt", "di", "dj", "wi", "wj", "z")}
        res = _lib.XiResult()
        res.num_pairs = cnt.ctypes.data
        for k, v in sums.items():
            setattr(res, k, v.ctypes.data)
        opts = _lib.XiOpts()
        opts.precision = int(precision)
        opts.accum = int(accum)
        opts.block_size = int(block_size)
        opts.blocks_per_cu = int(blocks_per_cu)
        opts.flags = int(flags)
        opts.wg_per_slot = int(wg_per_slot)
        g = grid.to_c()
        dev = self.dcat.dev
        check(lib().tos_xi_run(dev.h, self.dcat.h, self.h, C.byref(g), C.byref(opts), C.byref(res)),
              dev.h, "tos_xi_run")
        extra = {k: sums[k] for k in ("di", "dj", "wi", "wj", "z")} if accum == _lib.ACC_ALL else {}
        return XiHistogram(cnt, sums["didj"], sums["wgt"], extra, int(res.num_pixel_pairs_used),
                           int(res.num_guard_rechecks), float(res.kernel_ms), int(res.num_pairs_evaluated))


def healpix_radec2pix(order, ra, dec):
    ra = np.ascontiguousarray(ra, dtype=np.float64)
    dec = np.ascontiguousarray(dec, dtype=np.float64)
    out = np.zeros(len(ra), np.int32)
    check(lib().tos_healpix_radec2pix(int(order), len(ra), ra.ctypes.data, dec.ctypes.data,
                                      out.ctypes.data), what="tos_healpix_radec2pix")
    return out


def covariance_check(cov):
    """(is_pos_def, log_det) of a symmetric covariance (tos_covariance_check:
    FP64 Cholesky on the host; log_det is NaN when not positive definite)."""
    cov = np.ascontiguousarray(cov, dtype=np.float64)
    if cov.ndim != 2 or cov.shape[0] != cov.shape[1]:
        raise ValueError("covariance_check: a square matrix is required")
    pd = C.c_int32(0)
    ld = C.c_double(0)
    check(lib().tos_covariance_check(cov.shape[0], cov.ctypes.data, C.byref(pd), C.byref(ld)),
          what="tos_covariance_check")
    return bool(pd.value), float(ld.value)
