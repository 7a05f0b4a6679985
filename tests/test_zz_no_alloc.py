"""tos_xi_run_noalloc (the allocation-free pair pass) against tos_xi_run,
through the C-ABI: two calls in a row on one device, the second on a
smaller catalogue, each equal to tos_xi_run on the same input (counts
bit-exact, sums within the precision's tolerance).  Outputs are device
arrays from hipMalloc of the library's own HIP runtime."""
import ctypes as C

import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from tos_amd._lib import lib
from conftest import SUM_RTOL, rel_err, signed_sum_err, xi_err

pytestmark = pytest.mark.gpu

RTOL = SUM_RTOL


def hip():
    """The HIP runtime the library itself is linked against (a second HIP
    runtime in the process, such as torch's, does not see this context's
    device state)."""
    from tos_amd import _lib
    name = "libamdhip64.so.7"
    for line in open("/proc/self/maps"):
        if name in line:
            return C.CDLL(line.split()[-1])
    raise RuntimeError(f"{name} is not loaded by {_lib.LIB_PATH}")


class DeviceArray:
    """Plain device memory from hipMalloc, freed on close."""

    def __init__(self, H, nbytes):
        self.H = H
        self.p = C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def fill(self, byte):
        assert self.H.hipMemset(self.p, C.c_int(byte), C.c_size_t(self.nbytes)) == 0

    def get(self, dtype, n):
        out = np.empty(n, dtype=dtype)
        assert self.H.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0   # D2H
        return out

    def close(self):
        if self.p:
            self.H.hipFree(self.p)
            self.p = None


def run_noalloc(dev, pairs, grid, prec):
    lib()   # the library (and its HIP runtime) is loaded first
    H = hip()
    nb = grid.nbins_total
    bufs = [DeviceArray(H, nb * 8) for _ in range(3)]
    try:
        for b in bufs:
            b.fill(0xFF)     # poison: every output element must be written
        assert H.hipDeviceSynchronize() == 0
        g = grid.to_c()
        st = lib().tos_xi_run_noalloc(dev.h, pairs.dcat.h, pairs.h, C.byref(g), int(prec), bufs[0].p,
                                      bufs[1].p, bufs[2].p, None)
        if st != 0:
            raise tos_amd.TosError(st, lib().tos_last_error(dev.h).decode())
        assert H.hipDeviceSynchronize() == 0
        return bufs[0].get(np.uint64, nb), bufs[1].get(np.float64, nb), bufs[2].get(np.float64, nb)
    finally:
        for b in bufs:
            b.close()


@pytest.mark.parametrize("gname,prec", [("cartesian", tos_amd.FP32C), ("polar", tos_amd.FP64),
                                        ("observing", tos_amd.FP32C)])
def test_noalloc_equals_xi_run_twice(dev, gname, prec):
    grid = {"cartesian": GridSpec.cartesian(), "polar": GridSpec.polar(), "observing": GridSpec.observing()}[gname]
    big = synth.lya_catalog(600, area_deg2=60.0, seed=6100)
    small = synth.lya_catalog(90, area_deg2=9.0, seed=6101)
    for cat in (big, small):
        print(f"case: noalloc {gname} precision {prec} on {cat.n_los} sightlines", flush=True)
        dcat = dev.upload(cat)
        pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=8)
        ref = pairs.run(grid, precision=prec)        # builds the work items (and theta bins)
        cnt, didj, wgt = run_noalloc(dev, pairs, grid, prec)
        assert np.array_equal(cnt, ref.num_pairs)
        assert signed_sum_err(didj, ref.didj, ref.wgt) <= RTOL
        assert rel_err(wgt, ref.wgt) <= RTOL
        assert int(cnt.sum()) == ref.num_pixel_pairs_used
        pairs.free()
        dcat.free()


def test_noalloc_refuses_unprepared_pairs_and_huge_grids(dev):
    cat = synth.lya_catalog(60, area_deg2=6.0, seed=6102)
    grid = GridSpec.cartesian()
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=8)
    with pytest.raises(tos_amd.TosError, match="work items are not built"):
        run_noalloc(dev, pairs, grid, tos_amd.FP32C)
    pairs.run(grid, precision=tos_amd.FP32C)
    with pytest.raises(tos_amd.TosError, match="no theta bins"):
        run_noalloc(dev, pairs, GridSpec.observing(), tos_amd.FP64)
    huge = GridSpec.cartesian("[0:200]*2000", "[0:200]*2000")
    with pytest.raises(tos_amd.TosError):
        run_noalloc(dev, pairs, huge, tos_amd.FP64)
    pairs.free()
    dcat.free()


def test_noalloc_captured_in_a_graph(dev):
    """The allocation-free pass captured on a stream in global capture mode
    (any allocation or synchronous call inside would end the capture with
    an error) and replayed three times, with eager passes in between; every
    replay equals tos_xi_run."""
    lib()
    H = hip()
    cat = synth.lya_catalog(400, area_deg2=40.0, seed=6103)
    grid = GridSpec.cartesian()
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=8)
    ref = pairs.run(grid, precision=tos_amd.FP32C)
    nb = grid.nbins_total
    bufs = [DeviceArray(H, nb * 8) for _ in range(3)]
    stream, graph, exe = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert H.hipStreamCreate(C.byref(stream)) == 0
    try:
        g = grid.to_c()
        assert H.hipStreamBeginCapture(stream, 0) == 0        # hipStreamCaptureModeGlobal
        st = lib().tos_xi_run_noalloc(dev.h, dcat.h, pairs.h, C.byref(g), int(tos_amd.FP32C), bufs[0].p,
                                      bufs[1].p, bufs[2].p, stream)
        assert H.hipStreamEndCapture(stream, C.byref(graph)) == 0
        assert st == 0, lib().tos_last_error(dev.h).decode()
        assert H.hipGraphInstantiate(C.byref(exe), graph, None, None, C.c_size_t(0)) == 0
        for rep in range(3):
            print(f"case: graph replay {rep}", flush=True)
            for b in bufs:
                b.fill(0xFF)
            assert H.hipDeviceSynchronize() == 0
            assert H.hipGraphLaunch(exe, stream) == 0
            assert H.hipStreamSynchronize(stream) == 0
            cnt = bufs[0].get(np.uint64, nb)
            didj = bufs[1].get(np.float64, nb)
            wgt = bufs[2].get(np.float64, nb)
            assert np.array_equal(cnt, ref.num_pairs)
            assert signed_sum_err(didj, ref.didj, ref.wgt) <= RTOL
            assert rel_err(wgt, ref.wgt) <= RTOL
            eager = pairs.run(grid, precision=tos_amd.FP32C)     # eager work between replays
            assert np.array_equal(eager.num_pairs, ref.num_pairs)
    finally:
        if exe.value:
            H.hipGraphExecDestroy(exe)
        if graph.value:
            H.hipGraphDestroy(graph)
        H.hipStreamDestroy(stream)
        for b in bufs:
            b.close()
        pairs.free()
        dcat.free()
