"""GPU parity at the sizes of the benchmark workloads, against the CPU
oracle (XiEstimator.cc:121-305 restated) on the same seeded inputs:

  * 10,000 quasars over 1000 deg^2 (seed 2000): the whole pass, Cartesian,
    Polar and Observing grids, FP64 and FP32C.
  * 160,000 quasars over 10,000 deg^2 (seed 3000): per-sky-bin histograms
    of every k-th sky-bin task (the oracle runs those tasks only), Cartesian
    and Polar, FP64 and FP32C.
  * 128 mock realisations (seeds 4000+i, 2000-40,000 quasars over
    2000 deg^2): the batched, graph-replayed launch equals the 128 single
    runs, and the smallest mock equals the oracle.
  * a served stream of 8 jobs (1000-100,000 quasars, log-uniform, over
    1000 deg^2, Cartesian/Polar x FP64/FP32C): served results equal the
    direct path, and the smallest job equals the oracle.

Counts are bit-exact; sums (and xi) within 1e-10 relative in both modes
(FP32C bins with FP32 separations and an FP64 guard-band re-check, and
accumulates the reference's own FP32 products in FP64), the signed sum
w1 w2 d1 d2 with the cancellation floor of conftest.signed_sum_err."""
import os

import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from conftest import SUM_RTOL, rel_err, signed_sum_err, xi_err

pytestmark = pytest.mark.gpu

RTOL = SUM_RTOL
NTHREADS = min(16, os.cpu_count() or 8)


def check_vs_oracle(num_pairs, didj, wgt, ref):
    assert np.array_equal(np.asarray(num_pairs).astype(np.int64), ref["num_pairs"])
    assert signed_sum_err(didj, ref["didj"], ref["wgt"]) <= RTOL
    assert rel_err(wgt, ref["wgt"]) <= RTOL
    assert xi_err(didj, wgt, ref["didj"], ref["wgt"]) <= 10 * RTOL


def same(h, ref):
    assert np.array_equal(h.num_pairs, ref.num_pairs)
    assert signed_sum_err(h.didj, ref.didj, ref.wgt) <= RTOL
    assert rel_err(h.wgt, ref.wgt) <= RTOL
    assert h.num_pixel_pairs_used == ref.num_pixel_pairs_used


GRIDS = {"cartesian": GridSpec.cartesian(), "polar": GridSpec.polar(), "observing": GridSpec.observing()}


# ---------------------------------------------------------------------------
# 10,000 quasars over 1000 deg^2
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def qso10k():
    return synth.lya_catalog(10000, area_deg2=1000.0, seed=2000)


@pytest.mark.parametrize("gname", ["cartesian", "polar", "observing"])
def test_10k_quasars_full_pass_vs_oracle(dev, ora, qso10k, gname):
    grid = GRIDS[gname]
    print(f"case: 10,000 QSOs / 1000 deg2, {gname}: oracle", flush=True)
    ref = ora.xi_run(qso10k, grid, qso10k.scale, order=4, nthreads=NTHREADS)
    dcat = dev.upload(qso10k)
    pairs = dcat.pairs(grid.max_angular_scale(qso10k.scale), order=4, nthreads=NTHREADS, on_device=True)
    st = pairs.stats()
    assert st["num_sightline_pairs_used"] == ref["stats"]["num_sightline_pairs_used"]
    assert st["num_pixel_pairs"] == ref["stats"]["num_pixel_pairs"]
    for prec in (tos_amd.FP64, tos_amd.FP32C):
        print(f"case: 10,000 QSOs / 1000 deg2, {gname}, precision {prec}", flush=True)
        h = pairs.run(grid, precision=prec)
        assert h.num_pixel_pairs_used == ref["stats"]["num_pixel_pairs_used"]
        check_vs_oracle(h.num_pairs, h.didj, h.wgt, ref["raw"])
    pairs.free()
    dcat.free()


# ---------------------------------------------------------------------------
# 160,000 quasars over 10,000 deg^2: sampled sky-bin tasks
# ---------------------------------------------------------------------------
TASK_STRIDE = 24


@pytest.fixture(scope="module")
def qso160k():
    return synth.lya_catalog(160000, area_deg2=10000.0, seed=3000)


@pytest.mark.parametrize("gname", ["cartesian", "polar"])
def test_160k_quasars_sampled_skybin_tasks_vs_oracle(dev, ora, qso160k, gname):
    """Every 24th sky-bin task (first-insertion order, XiEstimator.cc:37-42)
    on the oracle against the same sky bins' subsample histograms of one
    per-sky-bin GPU pass, both precisions."""
    grid = GRIDS[gname]
    print(f"case: 160,000 QSOs / 10,000 deg2, {gname}: oracle on every "
          f"{TASK_STRIDE}th sky-bin task", flush=True)
    ref = ora.xi_run(qso160k, grid, qso160k.scale, order=4, nthreads=NTHREADS, task_stride=TASK_STRIDE,
                     task_phase=0, per_skybin=True)
    keys_ref = np.asarray(ref["skybin_keys"])
    assert len(keys_ref) >= 30
    dcat = dev.upload(qso160k)
    pairs = dcat.pairs(grid.max_angular_scale(qso160k.scale), order=4, nthreads=NTHREADS, on_device=True)
    for prec in (tos_amd.FP64, tos_amd.FP32C):
        print(f"case: 160,000 QSOs, {gname}, per-sky-bin pass, precision {prec}", flush=True)
        tot, keys, sn, sd, sw = pairs.run_per_skybin(grid, precision=prec)
        idx = np.searchsorted(keys, keys_ref)
        assert np.array_equal(keys[idx], keys_ref)
        for k, r in zip(idx, ref["per_skybin"]):
            check_vs_oracle(sn[k], sd[k], sw[k], r)
        used = sum(int(r["num_pairs"].sum()) for r in ref["per_skybin"])
        assert used == ref["stats"]["num_pixel_pairs_used"]
        assert int(sn[idx].sum()) == used
    pairs.free()
    dcat.free()


# ---------------------------------------------------------------------------
# 128 batched mocks
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def mocks(dev):
    sizes = synth.mock_sizes(128, 2000, 40000, seed=4000)
    grid = GridSpec.cartesian()
    out = []
    for i, n in enumerate(sizes):
        cat = synth.lya_catalog(int(n), area_deg2=2000.0, seed=4000 + i)
        dcat = dev.upload(cat)
        pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=NTHREADS, on_device=True)
        out.append((int(n), dcat, pairs))
        del cat
    yield out
    for _, dcat, pairs in out:
        pairs.free()
        dcat.free()


@pytest.mark.timeout(600)
def test_batch_of_128_mocks_equals_single_runs(dev, ora, mocks):
    grid = GridSpec.cartesian()
    print(f"case: {len(mocks)} mocks, {sum(n for n, _, _ in mocks)} QSOs: batched launch", flush=True)
    batch = [(p, grid, tos_amd.FP32C) for _, _, p in mocks]
    res, ms = dev.batch_run(batch, use_graph=True)
    res2, _ = dev.batch_run(batch, use_graph=True)     # a replay of the captured graph
    print(f"batched pass {ms:.1f} ms; single runs", flush=True)
    for (n, _, p), h, h2 in zip(mocks, res, res2):
        ref = p.run(grid, precision=tos_amd.FP32C)
        same(h, ref)
        same(h2, ref)
    # the smallest mock against the oracle
    k = int(np.argmin([n for n, _, _ in mocks]))
    cat = synth.lya_catalog(mocks[k][0], area_deg2=2000.0, seed=4000 + k)
    print(f"case: mock {k} ({mocks[k][0]} QSOs) against the oracle", flush=True)
    ref = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=NTHREADS)
    check_vs_oracle(res[k].num_pairs, res[k].didj, res[k].wgt, ref["raw"])


# ---------------------------------------------------------------------------
# Served stream in the 1000-100,000 quasar range
# ---------------------------------------------------------------------------
@pytest.mark.timeout(600)
def test_served_stream_of_8_jobs_equals_direct_and_oracle(dev, ora):
    from tos_amd.server import Server
    sizes = synth.job_sizes(8, 1000, 100000, seed=5000)
    rng = np.random.default_rng(5001)
    grids = [GridSpec.cartesian(), GridSpec.polar()]
    precs = [tos_amd.FP64, tos_amd.FP32C]
    jobs = []
    for k, n in enumerate(sizes):
        cat = synth.lya_catalog(int(n), area_deg2=1000.0, seed=5100 + k)
        jobs.append((cat, grids[int(rng.integers(2))], precs[int(rng.integers(2))]))
    assert {g.type for _, g, _ in jobs} == {tos_amd.CARTESIAN, tos_amd.POLAR}
    assert {p for _, _, p in jobs} == {tos_amd.FP64, tos_amd.FP32C}
    direct = []
    for cat, grid, prec in jobs:
        print(f"case: direct job {cat.n_los} QSOs type {grid.type} precision {prec}", flush=True)
        dcat = dev.upload(cat)
        pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=NTHREADS, on_device=True)
        direct.append(pairs.run(grid, precision=prec))
        pairs.free()
        dcat.free()
    print("case: the 8 jobs served together", flush=True)
    with Server(dev, nthreads=NTHREADS, search_on_device=True) as srv:
        ids = [srv.submit(cat, grid, prec) for cat, grid, prec in jobs]
        for jid, ref in zip(ids, direct):
            assert srv.wait(jid, 300000) == 0
            st, h, _ = srv.result(jid)
            assert st == 0
            same(h, ref)
    k = int(np.argmin(sizes))
    cat, grid, prec = jobs[k]
    print(f"case: served job {k} ({cat.n_los} QSOs) against the oracle", flush=True)
    ref = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=NTHREADS)
    check_vs_oracle(direct[k].num_pairs, direct[k].didj, direct[k].wgt, ref["raw"])
