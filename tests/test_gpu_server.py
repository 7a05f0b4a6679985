"""Served job stream (tos_server_* / tos_job_*): jobs of different sizes,
grids and precisions submitted while others run are admitted by HBM
footprint, cut into slices of bounded device time and batched onto
concurrent launches; every job's result must be the same alone and in the
mix (counts bit-exact, sums within 1e-10)."""
import time

import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from tos_amd.server import Server
from conftest import SUM_RTOL, rel_err, signed_sum_err, xi_err

pytestmark = pytest.mark.gpu

RTOL = SUM_RTOL


def same(h, ref):
    assert np.array_equal(h.num_pairs, ref.num_pairs)
    assert signed_sum_err(h.didj, ref.didj, ref.wgt) <= RTOL
    assert rel_err(h.wgt, ref.wgt) <= RTOL
    assert h.num_pixel_pairs_used == ref.num_pixel_pairs_used


GRIDS = [GridSpec.cartesian(), GridSpec.polar()]
PRECS = [tos_amd.FP64, tos_amd.FP32C]


@pytest.fixture(scope="module")
def stream():
    """Ten requests: sizes log-uniform over a decade, grid and precision
    drawn independently (the served-stream mix at test scale)."""
    rng = np.random.default_rng(5000)
    jobs = []
    for k in range(10):
        n = int(np.exp(rng.uniform(np.log(40), np.log(400))))
        cat = synth.lya_catalog(n, area_deg2=n / 10.0, seed=5000 + k)
        jobs.append((cat, GRIDS[rng.integers(2)], PRECS[rng.integers(2)]))
    return jobs


@pytest.fixture(scope="module")
def alone(dev, stream):
    """Each job through the direct path, one at a time."""
    out = []
    for cat, grid, prec in stream:
        dcat = dev.upload(cat)
        pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=8)
        out.append(pairs.run(grid, precision=prec))
        pairs.free()
        dcat.free()
    return out


@pytest.mark.parametrize("on_device", [False, True])
def test_served_alone_equals_direct(dev, stream, alone, on_device):
    print(f"served alone, search_on_device={on_device}", flush=True)
    with Server(dev, nthreads=8, search_on_device=on_device) as srv:
        for (cat, grid, prec), ref in zip(stream, alone):
            jid = srv.submit(cat, grid, prec)
            assert srv.wait(jid, 60000) == 0
            st, h, t = srv.result(jid)
            assert st == 0
            same(h, ref)
            assert t["submit_ms"] <= t["admit_ms"] <= t["start_ms"] <= t["done_ms"]
            assert t["footprint_bytes"] > 0


@pytest.mark.parametrize("on_device", [False, True])
def test_mix_equals_alone(dev, stream, alone, on_device):
    print(f"mix, search_on_device={on_device}", flush=True)
    rng = np.random.default_rng(1)
    with Server(dev, nthreads=8, search_on_device=on_device) as srv:
        ids = []
        for cat, grid, prec in stream:
            ids.append(srv.submit(cat, grid, prec))
            time.sleep(float(rng.uniform(0, 0.01)))
        for jid, ref in zip(ids, alone):
            assert srv.wait(jid, 60000) == 0
            st, h, t = srv.result(jid)
            assert st == 0
            same(h, ref)


def test_late_arrivals_start_while_a_large_job_runs(dev):
    """Continuous batching: small jobs submitted after a large one has been
    launched start on another launch slot before the large one is done."""
    big = synth.lya_catalog(40000, area_deg2=1000.0, seed=2000)
    small = [synth.lya_catalog(60, area_deg2=6.0, seed=7000 + k) for k in range(4)]
    with Server(dev, nthreads=8) as srv:
        jb = srv.submit(big, GridSpec.cartesian(), tos_amd.FP64)
        srv.wait_started(jb)     # the large job is searched, resident and launched
        js = [srv.submit(c, GridSpec.cartesian(), tos_amd.FP32C) for c in small]
        assert srv.wait(jb, 60000) == 0
        _, hb, tb = srv.result(jb)
        starts = []
        for j in js:
            assert srv.wait(j, 60000) == 0
            starts.append(srv.result(j)[2]["start_ms"])
        print("large job", tb, "small starts", starts, flush=True)
        assert tb["start_ms"] < min(starts) < tb["done_ms"]


def test_comparable_jobs_share_the_device_in_slices(dev):
    """Two large jobs, the second submitted once the first is running: their
    work items run in bounded slices, so the second starts before the first
    is done and both finish with the result they have alone."""
    a = synth.lya_catalog(30000, area_deg2=1000.0, seed=2100)
    b = synth.lya_catalog(30000, area_deg2=1000.0, seed=2101)
    grid = GridSpec.cartesian()
    ref = []
    for cat in (a, b):
        dcat = dev.upload(cat)
        pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=8)
        ref.append(pairs.run(grid, precision=tos_amd.FP64))
        pairs.free()
        dcat.free()
    with Server(dev, nthreads=8) as srv:
        ja = srv.submit(a, grid, tos_amd.FP64)
        srv.wait_started(ja)     # the first job is searched, resident and launched
        jb = srv.submit(b, grid, tos_amd.FP64)
        assert srv.wait(ja, 60000) == 0 and srv.wait(jb, 60000) == 0
        _, ha, ta = srv.result(ja)
        _, hb, tb = srv.result(jb)
    print("first", ta, "second", tb, flush=True)
    same(ha, ref[0])
    same(hb, ref[1])
    assert ta["start_ms"] < tb["start_ms"] < ta["done_ms"]


@pytest.mark.parametrize("on_device", [False, True])
def test_admission_by_footprint(dev, stream, alone, on_device):
    # footprints from a first pass, then a budget that fits one job at a time
    print(f"admission, search_on_device={on_device}", flush=True)
    fps = []
    with Server(dev, nthreads=8, search_on_device=on_device) as srv:
        for cat, grid, prec in stream[:4]:
            jid = srv.submit(cat, grid, prec)
            srv.wait(jid, 60000)
            fps.append(srv.result(jid)[2]["footprint_bytes"])
    budget = max(fps) + 1
    with Server(dev, hbm_budget_bytes=budget, max_inflight=2, nthreads=8, search_on_device=on_device) as srv:
        ids = [srv.submit(cat, grid, prec) for cat, grid, prec in stream[:4]]
        too_big = srv.submit(synth.lya_catalog(2000, area_deg2=200.0, seed=9), GridSpec.cartesian(), tos_amd.FP64)
        for jid, ref in zip(ids, alone[:4]):
            assert srv.wait(jid, 60000) == 0
            st, h, _ = srv.result(jid)
            same(h, ref)
        assert srv.wait(too_big, 60000) == 2     # TOS_ERR_NOMEM
        st, h, t = srv.result(too_big)
        assert st == 2 and h is None


def test_cancel_and_stop(dev, stream, alone):
    srv = Server(dev, nthreads=8)
    cat, grid, prec = stream[0]
    a = srv.submit(cat, grid, prec)
    srv.cancel(a)
    st = srv.wait(a, 60000)
    assert st in (0, 6)     # finished before the cancel, or cancelled
    srv.result(a)
    b = srv.submit(*stream[1])
    assert srv.wait(b, 60000) == 0
    same(srv.result(b)[1], alone[1])
    srv.stop()


def test_device_search_plates_and_footprint(dev, stream):
    """Plate sky bins through the served device search: the same result as
    the direct path, and the footprint admitted is the exact one (the
    device search counts the pairs before admission)."""
    cat = stream[0][0]
    grid = GridSpec.cartesian()
    plates = synth.plate_tiling(cat, spacing_deg=0.5)
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), plates=plates, nthreads=8)
    ref = pairs.run(grid, precision=tos_amd.FP64)
    pairs.free()
    dcat.free()
    res = {}
    for on_device in (False, True):
        with Server(dev, nthreads=8, search_on_device=on_device) as srv:
            jid = srv.submit(cat, grid, tos_amd.FP64, plates=plates)
            assert srv.wait(jid, 60000) == 0
            st, h, t = srv.result(jid)
        assert st == 0
        same(h, ref)
        res[on_device] = t
    assert res[True]["num_pixel_pairs"] == res[False]["num_pixel_pairs"]
    # the device search's tables are counted on top of the host figure
    assert res[True]["footprint_bytes"] >= res[False]["footprint_bytes"]


def test_plate_on_annulus_edge_goes_to_the_host_search(dev):
    """A plate centre exactly 1.5 deg + max_ang from a sightline: the
    device's cos test (PlateBins.h:78-81, |acos(c) - 1.5 deg| < max_ang)
    falls within rounding of its bound, so the device search leaves the
    decision to the host search (acos itself), both in tos_pairs_build and
    in the served job, which is sent back to a host worker.  Results equal
    the host search's."""
    cat = synth.lya_catalog(120, area_deg2=12.0, seed=2300)
    grid = GridSpec.cartesian()
    max_ang = grid.max_angular_scale(cat.scale)
    pol, pid, pra, pdec = synth.plate_tiling(cat, spacing_deg=2.5)
    a0 = np.arctan(1.0) * 4 / 180.0 * 1.5
    edge_id = int(pid.max()) + 1
    pid = np.append(pid, np.int32(edge_id)).astype(np.int32)
    pra = np.append(pra, cat.ra[0])
    pdec = np.append(pdec, cat.dec[0] + a0 + max_ang)
    pol = pol.copy()
    pol[1] = edge_id                      # the edge plate is occupied
    plates = (pol, pid, pra, pdec)
    # the precondition, on the host: sightline 0 against the edge plate
    # (types.h:37-39 operation order) sits within the search's 1e-12 margin
    # of cos(1.5 deg + max_ang)
    sd1, cd1, sr1, cr1 = np.sin(cat.dec[0]), np.cos(cat.dec[0]), np.sin(cat.ra[0]), np.cos(cat.ra[0])
    sd2, cd2, sr2, cr2 = np.sin(pdec[-1]), np.cos(pdec[-1]), np.sin(pra[-1]), np.cos(pra[-1])
    c = sd1 * sd2 + cd1 * cd2 * (sr1 * sr2 + cr1 * cr2)
    assert abs(c - np.cos(a0 + max_ang)) < 1e-13
    dcat = dev.upload(cat)
    host = dcat.pairs(max_ang, plates=plates, on_device=False)
    devp = dcat.pairs(max_ang, plates=plates, on_device=True)
    assert devp.stats() == host.stats()
    for x, y in zip(devp.download(), host.download()):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    ref = host.run(grid, precision=tos_amd.FP64)
    host.free()
    devp.free()
    dcat.free()
    with Server(dev, nthreads=8, search_on_device=True) as srv:
        jid = srv.submit(cat, grid, tos_amd.FP64, plates=plates)
        assert srv.wait(jid, 60000) == 0
        st, h, t = srv.result(jid)
    assert st == 0
    same(h, ref)


@pytest.mark.parametrize("on_device", [False, True])
def test_jobs_without_pairs(dev, on_device):
    """Edge cases served beside a normal job: one sightline (no pairs), two
    sightlines too far apart to pair, and an empty catalogue; all finish
    with empty histograms, and the normal job's result is unaffected."""
    grid = GridSpec.cartesian()
    one = synth.lya_catalog(1, area_deg2=1.0, seed=8100)
    apart = synth.lya_catalog(2, area_deg2=1.0, seed=8101)
    apart.ra[1] = apart.ra[0] + 1.0                 # 57 deg away
    empty = synth.lya_catalog(1, area_deg2=1.0, seed=8102).subset(np.zeros(0, dtype=np.int64))
    normal = synth.lya_catalog(80, area_deg2=8.0, seed=8103)
    dcat = dev.upload(normal)
    pairs = dcat.pairs(grid.max_angular_scale(normal.scale), order=4, nthreads=8)
    ref = pairs.run(grid, precision=tos_amd.FP32C)
    pairs.free()
    dcat.free()
    print(f"jobs without pairs, search_on_device={on_device}", flush=True)
    with Server(dev, nthreads=8, search_on_device=on_device) as srv:
        ids = [srv.submit(c, grid, tos_amd.FP32C, scale=normal.scale) for c in (one, apart, empty)]
        jn = srv.submit(normal, grid, tos_amd.FP32C)
        for jid in ids:
            assert srv.wait(jid, 60000) == 0
            st, h, t = srv.result(jid)
            assert st == 0
            assert int(h.num_pairs.sum()) == 0 and h.num_pixel_pairs_used == 0
            assert not h.wgt.any()
        assert srv.wait(jn, 60000) == 0
        st, h, _ = srv.result(jn)
    same(h, ref)


def test_oversize_grid_refused_at_submit(dev, stream, alone):
    """A grid whose LDS histogram cannot fit a served launch is refused by
    tos_job_submit (it would otherwise fail the batch it joins); jobs
    submitted around it finish as alone."""
    huge = GridSpec.cartesian("[0:200]*200", "[0:200]*200")
    with Server(dev, nthreads=8) as srv:
        a = srv.submit(*stream[0])
        with pytest.raises(tos_amd.TosError):
            srv.submit(stream[1][0], huge, tos_amd.FP64)
        b = srv.submit(*stream[2])
        for jid, ref in ((a, alone[0]), (b, alone[2])):
            assert srv.wait(jid, 60000) == 0
            same(srv.result(jid)[1], ref)


def test_result_while_another_thread_waits(dev):
    """tos_job_result from one thread while others block in tos_job_wait on
    the same job: the waiters return normally (the job is freed by the last
    of them); and tos_server_stop with a waiter blocked on a running job
    lets the job finish and the waiter leave before the server goes.  The
    job is large (hundreds of ms on the device), so the waiter threads are
    inside tos_job_wait long before it ends; one that arrives after the
    result was taken finds the id unknown (TOS_ERR_INVALID)."""
    import threading
    big = synth.lya_catalog(30000, area_deg2=1000.0, seed=2200)
    with Server(dev, nthreads=8) as srv:
        jid = srv.submit(big, GridSpec.cartesian(), tos_amd.FP64)
        srv.wait_started(jid)
        got = []
        ths = [threading.Thread(target=lambda: got.append(srv.wait(jid, 60000))) for _ in range(4)]
        for t in ths:
            t.start()
        assert srv.wait(jid, 60000) == 0
        st, h, _ = srv.result(jid)
        for t in ths:
            t.join()
        assert st == 0 and len(got) == 4 and all(g in (0, 1) for g in got)
    srv2 = Server(dev, nthreads=8)
    j2 = srv2.submit(big, GridSpec.cartesian(), tos_amd.FP64)
    srv2.wait_started(j2)
    out = []
    th = threading.Thread(target=lambda: out.append(srv2.wait(j2, 60000)))
    th.start()
    srv2.stop()
    th.join()
    assert out == [0]
