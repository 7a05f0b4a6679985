"""GPU parity: the HIP estimator against the CPU oracle on the same seeded
inputs.  Pair counts must be bit-exact; FP64 sums within 1e-10 relative
(FP64 mode) and the same for FP32C (its per-pair terms are the reference's
FP32 products, summed in FP64)."""
import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from conftest import SUM_RTOL, rel_err, signed_sum_err, xi_err

pytestmark = pytest.mark.gpu


def check_hist(h, ref, accum=tos_amd.ACC_CORE, rtol=SUM_RTOL):
    assert np.array_equal(h.num_pairs.astype(np.int64), ref["num_pairs"])
    assert signed_sum_err(h.didj, ref["didj"], ref["wgt"]) <= rtol
    assert rel_err(h.wgt, ref["wgt"]) <= rtol
    if accum == tos_amd.ACC_ALL:
        # di, dj: sums of signed terms w d, floored by their weight sums
        assert signed_sum_err(h.extra["di"], ref["di"], ref["wi"]) <= rtol
        assert signed_sum_err(h.extra["dj"], ref["dj"], ref["wj"]) <= rtol
        for k in ("wi", "wj", "z"):
            assert rel_err(h.extra[k], ref[k]) <= rtol, k
    # xi = didj / wgt (XiBin::finalize)
    assert xi_err(h.didj, h.wgt, ref["didj"], ref["wgt"]) <= 10 * rtol


@pytest.fixture(scope="module")
def mini():
    # about 1/30 of the 10,000-quasar configuration, same density
    return synth.lya_catalog(330, area_deg2=33.0, seed=2001)


def test_device_info(dev):
    info = dev.info()
    print(info)
    assert info["wavefront_size"] == 64
    assert "gfx950" in info["arch"]


def test_f64_arithmetic_matches_host(dev, ora):
    # sqrt: the pair kernels' own (no rescaling below 2^-767, a range no
    # squared separation from FP32 distances reaches), against the host's
    # correctly rounded sqrt over its whole domain: exponents from 2^-767 up,
    # both ends of each binade, +-0, +inf, NaN and negatives
    rng = np.random.default_rng(5)
    wide = 10.0 ** rng.uniform(np.log10(2.0 ** -767), 300, 20000)
    a = np.concatenate([rng.uniform(0, 5e4, 20000), rng.uniform(0, 1, 2000) ** 4 * 16, wide,
                        np.nextafter(2.0 ** np.arange(-760, 1000, 7), 0), 2.0 ** np.arange(-767, 1000, 7),
                        np.array([0.0, -0.0, 1.0, 2.0, 16.0, 2.0 ** -767, 4e4, -1.0, -0.5, -1.5, -4.9,
                                  2.0 ** 32 + 5, np.nan, np.inf, -np.inf, 1e19, -1e19])])
    b = np.concatenate([rng.uniform(1e-3, 5e3, len(a) - 5), np.array([3.0, 7.0, 0.1, 1e-5, 11.0])])
    s, q, c = dev.selftest_f64(a, b)
    with np.errstate(invalid="ignore", divide="ignore"):
        s_ref = np.sqrt(a)
        q_ref = a / b
    dom = (a >= 0) | np.isposinf(a)
    assert np.array_equal(s.view(np.uint64)[dom], s_ref.view(np.uint64)[dom])
    assert np.isnan(s[(a < 0) | np.isnan(a)]).all()
    fin = np.isfinite(q_ref)
    assert np.array_equal(q.view(np.uint64)[fin], q_ref.view(np.uint64)[fin])
    c_ref = np.array([ora.cvt_u32(float(x)) for x in a], dtype=np.uint32)
    assert np.array_equal(c, c_ref)


@pytest.mark.parametrize("prec", [tos_amd.FP64, tos_amd.FP32C])
@pytest.mark.parametrize("gridname", ["cartesian", "polar", "observing"])
def test_xi_parity_mini(dev, ora, mini, gridname, prec):
    grid = {"cartesian": GridSpec.cartesian(), "polar": GridSpec.polar(),
            "observing": GridSpec.observing()}[gridname]
    print(f"case: mini catalog {mini.n_los} LOS {mini.n_pix} px grid={gridname} prec={prec}", flush=True)
    ref = ora.xi_run(mini, grid, mini.scale, order=4, nthreads=8)
    dcat = dev.upload(mini)
    pairs = dcat.pairs(grid.max_angular_scale(mini.scale), order=4)
    st = pairs.stats()
    assert st["num_sightline_pairs_used"] == ref["stats"]["num_sightline_pairs_used"]
    assert st["num_pixel_pairs"] == ref["stats"]["num_pixel_pairs"]
    assert st["num_pixels"] == ref["stats"]["num_pixels"]
    h = pairs.run(grid, precision=prec)
    assert h.num_pixel_pairs_used == ref["stats"]["num_pixel_pairs_used"]
    check_hist(h, ref["raw"])
    if prec == tos_amd.FP32C:
        print(f"guard re-checks: {h.num_guard_rechecks} of {h.num_pixel_pairs_used}")
    pairs.free()
    dcat.free()


@pytest.mark.parametrize("flags", [tos_amd.XI_NO_WINDOW, tos_amd.XI_NO_TRANSPOSE,
                                   tos_amd.XI_NO_WINDOW | tos_amd.XI_NO_TRANSPOSE, tos_amd.XI_NO_INTERLEAVE,
                                   tos_amd.XI_NO_WINDOW | tos_amd.XI_NO_TRANSPOSE | tos_amd.XI_NO_INTERLEAVE])
@pytest.mark.parametrize("prec", [tos_amd.FP64, tos_amd.FP32C])
def test_xi_work_schemes_agree(dev, ora, mini, prec, flags):
    """The r_par window and the transposed remainders only change which
    wavefront evaluates a pair, never the result."""
    grid = GridSpec.polar() if flags == tos_amd.XI_NO_TRANSPOSE else GridSpec.cartesian()
    ref = ora.xi_run(mini, grid, mini.scale, order=4, nthreads=8)
    dcat = dev.upload(mini)
    pairs = dcat.pairs(grid.max_angular_scale(mini.scale), order=4)
    h = pairs.run(grid, precision=prec, flags=flags)
    check_hist(h, ref["raw"])
    full = pairs.run(grid, precision=prec)
    check_hist(full, ref["raw"])
    if flags & tos_amd.XI_NO_WINDOW:
        assert h.num_pairs_evaluated == pairs.stats()["num_pixel_pairs"]
        # the window skips pixel pairs the r_par cut rejects anyway
        assert full.num_pairs_evaluated < 0.95 * h.num_pairs_evaluated
    pairs.free()
    dcat.free()


@pytest.mark.parametrize("block,bpc,wps", [(256, 0, 1), (512, 2, 4), (1024, 1, 64), (768, 0, 0), (64, 1, 3)])
def test_launch_shape_keeps_results(dev, ora, mini, block, bpc, wps):
    """Workgroup size, workgroups per CU and workgroups per slot change only
    how the items are shared out: counts exact, sums within tolerance."""
    grid = GridSpec.cartesian()
    ref = ora.xi_run(mini, grid, mini.scale, order=4, nthreads=8)
    dcat = dev.upload(mini)
    pairs = dcat.pairs(grid.max_angular_scale(mini.scale), order=4)
    for prec in (tos_amd.FP64, tos_amd.FP32C):
        h = pairs.run(grid, precision=prec, block_size=block, blocks_per_cu=bpc, wg_per_slot=wps)
        check_hist(h, ref["raw"])
    pairs.free()
    dcat.free()


def test_unsorted_pixels_disable_the_window(dev, ora, mini):
    """Shuffle the pixels inside every sightline: the window needs ascending
    distances, so the library falls back to evaluating every pair."""
    rng = np.random.default_rng(8)
    cat = mini.subset(np.arange(mini.n_los))
    for k in range(cat.n_los):
        a, b = cat.offsets[k], cat.offsets[k + 1]
        perm = a + rng.permutation(b - a)
        for f in ("value", "loglam", "weight", "distance"):
            arr = getattr(cat, f)
            arr[a:b] = arr[perm]
    grid = GridSpec.cartesian()
    ref = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=8)
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4)
    h = pairs.run(grid, precision=tos_amd.FP32C)
    check_hist(h, ref["raw"])
    st = pairs.stats()
    assert h.num_pairs_evaluated == st["num_pixel_pairs"]
    pairs.free()
    dcat.free()


@pytest.mark.parametrize("prec", [tos_amd.FP64, tos_amd.FP32C])
def test_xi_all_fields(dev, ora, mini, prec):
    grid = GridSpec.cartesian()
    ref = ora.xi_run(mini, grid, mini.scale, order=4, nthreads=8)
    dcat = dev.upload(mini)
    pairs = dcat.pairs(grid.max_angular_scale(mini.scale), order=4)
    h = pairs.run(grid, precision=prec, accum=tos_amd.ACC_ALL)
    check_hist(h, ref["raw"], accum=tos_amd.ACC_ALL)
    pairs.free()
    dcat.free()


@pytest.mark.parametrize("mode", [tos_amd.BRUTE_FP64, tos_amd.BRUTE_FP32])
@pytest.mark.parametrize("n,mn,mx,nbins", [(3000, 0.0, 100.0, 50),     # power-of-two width (100 Mpc/h in 50 bins)
                                           (3001, 5.0, 95.0, 37),      # general width, nonzero minimum
                                           (2000, 0.0, 60.0, 1000),    # many bins: fewer histogram copies
                                           (65, 0.0, 100.0, 50)])      # one 64-point block and one lane
def test_brute3d_parity(dev, ora, mode, n, mn, mx, nbins):
    x, y, z, d, _ = synth.points3d(n, seed=1000 + n)
    w = np.random.default_rng(n).uniform(0.5, 2.0, n)
    if mode == tos_amd.BRUTE_FP64:
        ds, ws, cn = ora.brute3d_f64(x, y, z, d, w, mn, mx, nbins)
    else:
        f = [np.asarray(a, dtype=np.float32) for a in (x, y, z, d, w)]
        x, y, z, d, w = [a.astype(np.float64) for a in f]
        ds, ws, cn = ora.brute3d_f32(*f, mn, mx, nbins)
    print(f"brute3d mode={mode} n={n} range=[{mn},{mx}) nbins={nbins}", flush=True)
    gds, gws, gcn, ms = dev.brute3d(x, y, z, d, w, mn, mx, nbins, mode=mode)
    assert np.array_equal(gcn.astype(np.int64), cn)
    assert rel_err(gds, ds) <= SUM_RTOL
    assert rel_err(gws, ws) <= SUM_RTOL


@pytest.mark.parametrize("gridname", ["cartesian", "polar"])
@pytest.mark.parametrize("prec", [tos_amd.FP64, tos_amd.FP32C])
def test_per_skybin_subsamples(dev, ora, mini, gridname, prec):
    """Per-sky-bin histograms (XiEstimator.h:78, skybin_xis_) in one batched
    pass: every occupied sky bin, ascending keys, each equal to the oracle's
    task for that bin; the total is their sum (XiEstimator.cc:298-305)."""
    grid = getattr(GridSpec, gridname)()
    ref = ora.xi_run(mini, grid, mini.scale, order=4, nthreads=8, per_skybin=True)
    dcat = dev.upload(mini)
    pairs = dcat.pairs(grid.max_angular_scale(mini.scale), order=4)
    tot, keys, sn, sd, sw = pairs.run_per_skybin(grid, precision=prec)
    order = np.argsort(ref["skybin_keys"])
    assert np.array_equal(keys, np.asarray(ref["skybin_keys"])[order])
    per = ref["per_skybin"][order]
    assert np.array_equal(sn.astype(np.int64), per["num_pairs"])
    for k in range(len(keys)):
        assert signed_sum_err(sd[k], per["didj"][k], per["wgt"][k]) <= SUM_RTOL
        assert rel_err(sw[k], per["wgt"][k]) <= SUM_RTOL
    check_hist(tot, ref["raw"])
    assert np.array_equal(sn.sum(0), tot.num_pairs)
    pairs.free()
    dcat.free()


@pytest.mark.parametrize("gname", ["cartesian", "polar", "observing"])
@pytest.mark.parametrize("prec", [tos_amd.FP64, tos_amd.FP32C])
def test_golden_c1mini(dev, gname, prec):
    """The device against the committed golden histograms of
    tests/golden/c1mini.npz (inputs stored in the file)."""
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_golden as mg
    from tos_amd.synth import Catalog
    z = np.load(os.path.join(here, "c1mini.npz"))
    cat = Catalog(z["ra"], z["dec"], z["offsets"], z["value"], z["loglam"], z["weight"], z["distance"])
    cat.scale = float(z["scale"])
    a1, a2 = mg.GRIDS[gname]
    grid = getattr(GridSpec, gname)(a1, a2)
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4)
    assert pairs.stats()["num_pixel_pairs"] == int(z[f"{gname}/stats/num_pixel_pairs"])
    h = pairs.run(grid, precision=prec, accum=tos_amd.ACC_ALL)
    ref = {f: z[f"{gname}/{f}"] for f in ("num_pairs", "didj", "wgt", "di", "dj", "wi", "wj", "z")}
    check_hist(h, ref, accum=tos_amd.ACC_ALL)
    pairs.free()
    dcat.free()


@pytest.mark.parametrize("mode", [tos_amd.BRUTE_FP64, tos_amd.BRUTE_FP32])
def test_golden_brute3d(dev, mode):
    """bruteGPU on the reference's sample files against the golden sums."""
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_golden as mg
    g = np.load(os.path.join(here, "brute3d.npz"))
    tag = "f64" if mode == tos_amd.BRUTE_FP64 else "f32"
    for name in mg.SAMPLES:
        cols = np.loadtxt(os.path.join(here, name + ".dat"), usecols=(0, 1, 2, 3, 4)).T
        ds, ws, cn, _ = dev.brute3d(*cols, *mg.C0_RANGE, mode=mode)
        assert np.array_equal(cn.astype(np.int64), g[f"{name}/{tag}/count"]), name
        assert rel_err(ds, g[f"{name}/{tag}/dsum"]) <= SUM_RTOL, name
        assert rel_err(ws, g[f"{name}/{tag}/wsum"]) <= SUM_RTOL, name


def test_catalog_upload_rejects_bad_input(dev):
    """tos_catalog_upload validates on the host before any device work: a
    non-finite pixel or decreasing offsets are errors, and when both occur
    the sightline that comes first decides which one is reported."""
    import dataclasses
    cat = synth.lya_catalog(60, area_deg2=6.0, seed=11)
    o = cat.offsets

    def with_nan(c, los):
        v = c.value.copy()
        v[o[los] + 1] = np.nan
        return dataclasses.replace(c, value=v)

    def with_bad_offsets(c, los):
        off = c.offsets.copy()
        off[los + 1] = off[los] - 1        # sightline `los` ends before it starts
        return dataclasses.replace(c, offsets=off)

    cases = [(with_nan(cat, 30), "non-finite"), (with_bad_offsets(cat, 20), "non-decreasing"),
             (with_bad_offsets(with_nan(cat, 5), 20), "non-finite"),
             (with_bad_offsets(with_nan(cat, 30), 20), "non-decreasing")]
    for bad, msg in cases:
        with pytest.raises(tos_amd.TosError, match=msg):
            dev.upload(bad)
    # the intact catalogue still uploads and runs
    d = dev.upload(cat)
    d.free()


@pytest.mark.parametrize("accum,grid", [
    (tos_amd.ACC_ALL, GridSpec.cartesian(z="[3.556:3.735]*2")),      # 50 rows of 101 slots, 60 B each
    (tos_amd.ACC_ALL, GridSpec.polar("[0:200]*50", "[0:1]*40", z="[3.556:3.735]*2")),
    (tos_amd.ACC_CORE, GridSpec.cartesian("[0:200]*200", "[0:200]*200")),   # 200 rows of 201 slots
], ids=["all-cart-50x50x2", "all-polar-50x40x2", "core-cart-200x200"])
@pytest.mark.parametrize("prec", [tos_amd.FP64, tos_amd.FP32C])
def test_grid_larger_than_one_lds_histogram(dev, ora, mini, accum, grid, prec):
    """Grids whose histogram does not fit one workgroup's LDS run in bands
    of axis-0 rows (one pass over the pairs per band): the result equals the
    oracle's, every XiBin field with TOS_ACC_ALL."""
    print(f"case: banded grid {grid.specs} accum {accum} precision {prec}", flush=True)
    ref = ora.xi_run(mini, grid, mini.scale, order=4, nthreads=8)
    dcat = dev.upload(mini)
    pairs = dcat.pairs(grid.max_angular_scale(mini.scale), order=4)
    h = pairs.run(grid, precision=prec, accum=accum)
    assert h.num_pixel_pairs_used == ref["stats"]["num_pixel_pairs_used"]
    check_hist(h, ref["raw"], accum=accum)
    pairs.free()
    dcat.free()
