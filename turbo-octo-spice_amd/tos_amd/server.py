"""Python front end of the job server (tos_server_* / tos_job_*): a served
stream of xi jobs with HBM-footprint admission and continuous batching."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .estimator import XiHistogram, sky_desc


class Server:
    def __init__(self, dev, hbm_budget_bytes: int = 64 << 30, max_inflight: int = 64, nthreads: int = 8,
                 search_on_device: bool = False):
        self.dev = dev
        o = _lib.ServerOpts()
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        o.max_inflight = int(max_inflight)
        o.nthreads = int(nthreads)
        o.search_on_device = int(bool(search_on_device))
        h = C.c_void_p()
        check(lib().tos_server_start(dev.h, C.byref(o), C.byref(h)), dev.h, "tos_server_start")
        self.h = h
        self.grids = {}

    def _desc(self, cat, grid, precision, scale, order, plates):
        s, keep = cat.to_c()
        sky, keep2 = sky_desc(order, plates)
        d = _lib.JobDesc()
        d.sightlines = s
        d.sky = sky
        d.grid = grid.to_c()
        d.scale = float(cat.scale if scale is None else scale)
        d.precision = int(precision)
        return d, keep + keep2

    def submit(self, cat, grid, precision=_lib.FP64, scale=None, order=4, plates=None) -> int:
        d, keep = self._desc(cat, grid, precision, scale, order, plates)
        jid = C.c_uint64(0)
        check(lib().tos_job_submit(self.h, C.byref(d), C.byref(jid)), self.dev.h, "tos_job_submit")
        self.grids[jid.value] = grid
        return int(jid.value)

    def footprint(self, cat, grid, precision=_lib.FP64, scale=None, order=4) -> int:
        d, keep = self._desc(cat, grid, precision, scale, order, None)
        b = C.c_int64(0)
        check(lib().tos_job_footprint(C.byref(d), C.byref(b)), what="tos_job_footprint")
        return int(b.value)

    def poll(self, jid: int) -> int:
        return int(lib().tos_job_poll(self.h, C.c_uint64(jid)))

    def wait(self, jid: int, timeout_ms: float = -1.0) -> int:
        return int(lib().tos_job_wait(self.h, C.c_uint64(jid), float(timeout_ms)))

    def timing(self, jid: int) -> dict:
        """The job's timing so far (tos_job_timing_get); start_ms is 0 until
        its first launch."""
        t = _lib.JobTiming()
        check(lib().tos_job_timing_get(self.h, C.c_uint64(jid), C.byref(t)), what="tos_job_timing_get")
        return {k: getattr(t, k) for k, _ in t._fields_}

    def wait_started(self, jid: int, timeout_s: float = 60.0) -> dict:
        """Poll until the job has been launched (or has finished)."""
        import time
        t_end = time.monotonic() + timeout_s
        while True:
            t = self.timing(jid)
            if t["start_ms"] > 0 or self.poll(jid) != 5:   # 5: TOS_ERR_NOT_READY
                return t
            if time.monotonic() > t_end:
                raise TimeoutError(f"job {jid} not launched within {timeout_s} s")
            time.sleep(0.002)

    def cancel(self, jid: int) -> int:
        return int(lib().tos_job_cancel(self.h, C.c_uint64(jid)))

    def result(self, jid: int):
        """(status, XiHistogram or None, timing dict); releases the job."""
        grid = self.grids.pop(jid)
        nb = grid.nbins_total
        cnt, didj, wgt = np.zeros(nb, np.uint64), np.zeros(nb), np.zeros(nb)
        r = _lib.XiResult()
        r.num_pairs, r.didj, r.wgt = cnt.ctypes.data, didj.ctypes.data, wgt.ctypes.data
        t = _lib.JobTiming()
        st = int(lib().tos_job_result(self.h, C.c_uint64(jid), C.byref(r), C.byref(t)))
        timing = {k: getattr(t, k) for k, _ in t._fields_}
        if st != _lib.TOS_OK:
            return st, None, timing
        h = XiHistogram(cnt, didj, wgt, {}, int(r.num_pixel_pairs_used), int(r.num_guard_rechecks),
                        float(r.kernel_ms), int(r.num_pairs_evaluated))
        return st, h, timing

    def stop(self):
        if self.h:
            check(lib().tos_server_stop(self.h), what="tos_server_stop")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()
