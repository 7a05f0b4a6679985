"""Batched launches (tos_batch_run): many jobs of different sizes, grids and
precisions in one launch must give each job the result of its own
tos_xi_run (counts bit-exact, sums within 1e-10), eagerly and through
hipGraph replays; the autotuned launch shape must not change results."""
import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from conftest import SUM_RTOL, rel_err, signed_sum_err, xi_err

pytestmark = pytest.mark.gpu

RTOL = SUM_RTOL


def same(h, ref, evaluated=True):
    assert np.array_equal(h.num_pairs, ref.num_pairs)
    assert signed_sum_err(h.didj, ref.didj, ref.wgt) <= RTOL
    assert rel_err(h.wgt, ref.wgt) <= RTOL
    assert h.num_pixel_pairs_used == ref.num_pixel_pairs_used
    assert h.num_guard_rechecks == ref.num_guard_rechecks
    if evaluated:
        assert h.num_pairs_evaluated == ref.num_pairs_evaluated


@pytest.fixture(scope="module")
def jobs(dev):
    """Uneven mocks (own seeds and sizes) with mixed grids and precisions,
    plus a one-sightline job with no pairs."""
    spec = [(60, 6.0, 11, GridSpec.cartesian(), tos_amd.FP32C),
            (260, 26.0, 12, GridSpec.cartesian(), tos_amd.FP32C),
            (140, 14.0, 13, GridSpec.cartesian(), tos_amd.FP64),
            (200, 20.0, 14, GridSpec.polar(), tos_amd.FP32C),
            (90, 9.0, 15, GridSpec.polar(), tos_amd.FP64),
            (120, 12.0, 16, GridSpec.observing(), tos_amd.FP64),
            (1, 1.0, 17, GridSpec.cartesian(), tos_amd.FP32C)]
    out = []
    for n, area, seed, grid, prec in spec:
        cat = synth.lya_catalog(n, area_deg2=area, seed=seed)
        dcat = dev.upload(cat)
        pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=8)
        out.append((cat, dcat, pairs, grid, prec))
    yield out
    for _, dcat, pairs, _, _ in out:
        pairs.free()
        dcat.free()


def test_batch_equals_single_runs(dev, jobs):
    print("batch of", len(jobs), "jobs", flush=True)
    singles = [pairs.run(grid, precision=prec) for _, _, pairs, grid, prec in jobs]
    res, ms = dev.batch_run([(p, g, pr) for _, _, p, g, pr in jobs])
    assert ms > 0
    for h, ref in zip(res, singles):
        same(h, ref)
    assert res[-1].num_pixel_pairs_used == 0


def test_batch_against_oracle(dev, ora, jobs):
    cat, _, pairs, grid, prec = jobs[1]
    res, _ = dev.batch_run([(pairs, grid, prec)])
    ref = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=8)["raw"]
    assert np.array_equal(res[0].num_pairs.astype(np.int64), ref["num_pairs"])
    assert signed_sum_err(res[0].didj, ref["didj"], ref["wgt"]) <= RTOL
    assert rel_err(res[0].wgt, ref["wgt"]) <= RTOL


def test_graph_replays_match_eager(dev, jobs):
    batch = [(p, g, pr) for _, _, p, g, pr in jobs]
    eager, _ = dev.batch_run(batch, use_graph=False)
    for rep in range(3):
        print("graph replay", rep, flush=True)
        res, ms = dev.batch_run(batch, use_graph=True)
        for h, ref in zip(res, eager):
            same(h, ref)
        # eager work on the same context between replays
        _, _, pairs, grid, prec = jobs[2]
        pairs.run(grid, precision=prec)


def test_changed_batch_is_prepared_again(dev, jobs):
    a = [(p, g, pr) for _, _, p, g, pr in jobs[:3]]
    b = [(p, g, pr) for _, _, p, g, pr in jobs[3:6]]
    ra, _ = dev.batch_run(a, use_graph=True)
    rb, _ = dev.batch_run(b, use_graph=True)
    ra2, _ = dev.batch_run(a, use_graph=True)
    for h, ref in zip(ra2, ra):
        same(h, ref)
    for (_, _, p, g, pr), h in zip(jobs[3:6], rb):
        same(h, p.run(g, precision=pr))


def test_autotune_keeps_results(dev, jobs):
    _, _, pairs, grid, prec = jobs[1]
    before = pairs.run(grid, precision=prec)
    block, bpc, ms = dev.autotune(pairs, grid, prec)
    print("tuned", block, bpc, ms, flush=True)
    assert block in (256, 512, 1024) and bpc in (1, 2, 4, 8) and ms > 0
    shape = dev.tuned_shape(grid, prec)
    print("tuned shape", shape, flush=True)
    assert shape["block"] == block and shape["blocks_per_cu"] == bpc
    assert shape["wg_per_slot"] in (0, 4, 8, 16, 64)
    assert shape["step_cap"] in (0, 1024, 2048, 8192, 16384)
    same(pairs.run(grid, precision=prec), before)
    res, _ = dev.batch_run([(pairs, grid, prec)])
    same(res[0], before)


def test_batch_rejects_bad_jobs(dev, jobs):
    _, _, pairs, grid, _ = jobs[0]
    with pytest.raises(tos_amd.TosError):
        dev.batch_run([(pairs, grid, 7)])
