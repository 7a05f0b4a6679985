"""GPU against the CPU oracle on seeded random grids: random axis minima
(zero, positive and slightly negative), bin widths (powers of two and
others), bin counts (including grids too large for one fixed LDS
histogram) and z axes, for the Cartesian, Polar and Observing bin
functions (XiEstimator.cc:168-294), both precisions.  The FP32 path's
guard band, its exact-zero and SIMPLE shortcuts and the window are
derived per grid, so every grid is a new set of constants.  All of them
also run as the jobs of one batched launch.
Counts bit-exact, sums within 1e-10 relative."""
import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from conftest import SUM_RTOL, rel_err, signed_sum_err

pytestmark = pytest.mark.gpu

N_CASES = 36


def axis(lo, width, n):
    return f"[{lo:.9g}:{lo + width * n:.9g}]*{n}"


def random_grid(k):
    rng = np.random.default_rng(9100 + k)
    kind = ("cartesian", "polar", "observing")[k % 3]
    nz = int(rng.integers(1, 4))
    z = axis(3.556, (3.735 - 3.556) / nz, nz)

    def width(scale):
        if rng.random() < 0.4:
            return scale * 2.0 ** int(rng.integers(-2, 3))
        return scale * float(rng.uniform(0.3, 3.0))

    def start(scale):
        u = rng.random()
        if u < 0.5:
            return 0.0
        if u < 0.8:
            return float(rng.uniform(0.0, 3.0)) * scale
        return -float(rng.uniform(0.0, 0.9)) * scale

    if kind == "cartesian":
        w0, w1 = width(4.0), width(4.0)
        n0, n1 = int(rng.integers(5, 60)), int(rng.integers(5, 60))
        return GridSpec.cartesian(axis(start(w0), w0, n0), axis(start(w1), w1, n1), z)
    if kind == "polar":
        w0 = width(4.0)
        n0 = int(rng.integers(5, 60))
        n1 = int(rng.integers(4, 40))
        mu0 = start(1.0 / n1) if rng.random() < 0.7 else 0.0
        w1 = (1.0 - mu0) / n1 if rng.random() < 0.6 else width(1.0 / n1)
        return GridSpec.polar(axis(start(w0), w0, n0), axis(mu0, w1, n1), z)
    w0 = width(1e-3)
    n0 = int(rng.integers(5, 30))
    n1 = int(rng.integers(2, 10))
    w1 = float(rng.uniform(0.004, 0.012))
    return GridSpec.observing(axis(start(w0), w0, n0), axis(0.0, w1, n1), z)


@pytest.fixture(scope="module")
def rcat():
    return synth.lya_catalog(120, area_deg2=12.0, seed=9099)


@pytest.fixture(scope="module")
def rpairs(dev, rcat):
    dcat = dev.upload(rcat)
    yield dcat
    dcat.free()


@pytest.mark.parametrize("k", range(N_CASES))
def test_random_grid_matches_oracle(dev, ora, rcat, rpairs, k):
    grid = random_grid(k)
    ref = ora.xi_run(rcat, grid, rcat.scale, order=4, nthreads=8)["raw"]
    pairs = rpairs.pairs(grid.max_angular_scale(rcat.scale), order=4, nthreads=8)
    try:
        for prec in (tos_amd.FP64, tos_amd.FP32C):
            print(f"case {k}: grid {grid.specs} precision {prec}", flush=True)
            h = pairs.run(grid, precision=prec)
            assert np.array_equal(h.num_pairs.astype(np.int64), ref["num_pairs"])
            assert signed_sum_err(h.didj, ref["didj"], ref["wgt"]) <= SUM_RTOL
            assert rel_err(h.wgt, ref["wgt"]) <= SUM_RTOL
    finally:
        pairs.free()


def test_random_grids_in_one_batch(dev, rcat, rpairs):
    """All seeded grids in both precisions as the jobs of one batched launch
    (one launch per kernel variant, each job's histogram from its own
    partial rows): every job equals its own single run."""
    lists = [rpairs.pairs(random_grid(k).max_angular_scale(rcat.scale), order=4, nthreads=8)
             for k in range(N_CASES)]
    try:
        jobs = [(lists[k], random_grid(k), prec) for k in range(N_CASES)
                for prec in (tos_amd.FP64, tos_amd.FP32C)]
        print(f"batch of {len(jobs)} jobs", flush=True)
        res, _ = dev.batch_run(jobs)
        for (pairs, grid, prec), h in zip(jobs, res):
            ref = pairs.run(grid, precision=prec)
            assert np.array_equal(h.num_pairs, ref.num_pairs), grid.specs
            assert signed_sum_err(h.didj, ref.didj, ref.wgt) <= SUM_RTOL
            assert rel_err(h.wgt, ref.wgt) <= SUM_RTOL
            assert h.num_pixel_pairs_used == ref.num_pixel_pairs_used
    finally:
        for p in lists:
            p.free()


def test_random_grids_cover_the_variants():
    """The seeded cases include zero and nonzero minima, power-of-two and
    other widths, several z bins and a grid larger than one fixed LDS
    histogram (2560 slots)."""
    grids = [random_grid(k) for k in range(N_CASES)]
    assert any(g.nbins_total > 2560 for g in grids)
    assert any(int(g.nbins[2]) > 1 for g in grids)
    mins = [float(s.split(":")[0][1:]) for g in grids for s in g.specs[:2]]
    assert any(m == 0.0 for m in mins) and any(m > 0 for m in mins) and any(m < 0 for m in mins)
