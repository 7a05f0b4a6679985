"""The reference's edge-case arithmetic on the GPU, both precisions, against
the CPU oracle (the same cases tests/test_oracle_pins.py pins on the oracle):

  * Polar NaN path (XiEstimator.cc:183-192): coincident sightlines give
    r^2 = f32(d1^2) - d1^2 + ..., negative for some pixels: sqrt gives NaN,
    the cast gives r bin 0, 'r == 0' is false, mu is NaN -> mu bin 0, so
    the pair lands in bin (0, 0, iz); r^2 == 0 is skipped by 'r == 0'.
  * the z axis: cast without a range check and scaled by 1/NBins
    (XiEstimator.cc:174, 232, 251-253): axes above the data, one that wraps
    (x86 cast of t <= -1), two z bins, and axes on which some sightline
    pairs share one z bin (the constant-z shortcut) and others do not, with
    the shortcut on and off.
  * negative coordinates on the r_par / r_perp / r / mu axes (E9: the x86
    cast maps (-1, 0) to bin 0 and t <= -1 to a rejected index).
  * the unchecked theta bin of the Observing grid (XiEstimator.cc:273,283):
    an axis starting above 0 (small angles wrap to a huge index, which the
    unsigned flat-index arithmetic carries back into range) and a pair list
    wider than the theta axis (theta bins past the axis carry into the next
    dloglam row; past the last row they are skipped).
Counts bit-exact, sums within 1e-10 relative."""
import math

import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from conftest import SUM_RTOL, rel_err, signed_sum_err, xi_err

pytestmark = pytest.mark.gpu

RTOL = SUM_RTOL


def check(h, ref):
    assert np.array_equal(h.num_pairs.astype(np.int64), ref["num_pairs"])
    assert signed_sum_err(h.didj, ref["didj"], ref["wgt"]) <= RTOL
    assert rel_err(h.wgt, ref["wgt"]) <= RTOL


def run_both(dev, ora, cat, grid, flags=(0,), order=4):
    ref = ora.xi_run(cat, grid, cat.scale, order=order, nthreads=8)
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=order, nthreads=8)
    out = {}
    for prec in (tos_amd.FP64, tos_amd.FP32C):
        for f in flags:
            print(f"case: {cat.n_los} LOS grid {grid.specs} precision {prec} flags {f}", flush=True)
            h = pairs.run(grid, precision=prec, flags=f)
            check(h, ref["raw"])
            assert h.num_pixel_pairs_used == ref["stats"]["num_pixel_pairs_used"]
            out[(prec, f)] = h
    pairs.free()
    dcat.free()
    return ref, out


def catalog(ra, dec, counts, d, l, w, v, scale):
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cat = synth.Catalog(np.asarray(ra, float), np.asarray(dec, float), off, np.asarray(v, np.float32),
                        np.asarray(l, np.float32), np.asarray(w, np.float32), np.asarray(d, np.float32))
    cat.scale = scale
    return cat


def float_with_e1(sign, start=4000.0):
    """A float distance whose float square is below (sign < 0), above
    (sign > 0) or equal to (sign == 0) its exact square."""
    d = np.float32(start)
    while True:
        e1 = float(np.float64(d * d)) - float(d) ** 2
        if (sign < 0 and e1 < 0) or (sign > 0 and e1 > 0) or (sign == 0 and e1 == 0):
            return float(d)
        d = np.nextafter(d, np.float32(5000.0))


# ---------------------------------------------------------------------------
# Polar NaN path
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("sign,expect_used", [(-1, 1), (1, 1), (0, 0)])
def test_polar_nan_path_hand_case(dev, ora, sign, expect_used):
    d = float_with_e1(sign) if sign != 0 else 4096.0
    cat = catalog([0.3, 0.3], [0.1, 0.1], [1, 1], [d, d], [3.6, 3.6], [1.0, 2.0], [0.5, 0.5], 3586.0)
    ref, out = run_both(dev, ora, cat, GridSpec.polar())
    assert ref["stats"]["num_pixel_pairs_used"] == expect_used
    for h in out.values():
        assert int(h.num_pairs[0]) == expect_used


def test_coincident_sightlines(dev, ora):
    """Six sightlines at one position, pixels on the common wavelength grid:
    every pixel pair at equal wavelengths has r_par = 0 and r^2 = e1, every
    other pair r ~ r_par, mu ~ 1 (at the mu = 1 edge), r_perp ~ 0."""
    base = synth.lya_catalog(6, area_deg2=1.0, seed=8300)
    base.ra[:] = base.ra[0]
    base.dec[:] = base.dec[0]
    for gname in ("polar", "cartesian"):
        grid = getattr(GridSpec, gname)()
        ref, out = run_both(dev, ora, base, grid, flags=(0, tos_amd.XI_NO_WINDOW))
        if gname == "polar":
            # the NaN path: pairs at equal distance with f32(d^2) < d^2
            assert int(ref["raw"]["num_pairs"][0]) > 0


# ---------------------------------------------------------------------------
# z axis
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def zcat():
    return synth.lya_catalog(150, area_deg2=15.0, seed=21)


@pytest.mark.parametrize("z", ["[3.8:4.0]*1", "[5:6]*1", "[3.556:3.735]*2", "[2.6:3.6]*1", "[1.6:3.7]*2"])
@pytest.mark.parametrize("gname", ["cartesian", "polar", "observing"])
def test_z_axis_quirks(dev, ora, zcat, gname, z):
    make = {"cartesian": lambda: GridSpec.cartesian(z=z), "polar": lambda: GridSpec.polar(z=z),
            "observing": lambda: GridSpec.observing(z=z)}[gname]
    ref, _ = run_both(dev, ora, zcat, make(), flags=(0, tos_amd.XI_NO_ZCONST))
    if z == "[5:6]*1":
        # the wrapped cast shifts every flat index down by one
        base = ora.xi_run(zcat, getattr(GridSpec, gname)(), zcat.scale, order=4, nthreads=8)["raw"]["num_pairs"]
        assert np.array_equal(ref["raw"]["num_pairs"][:-1], base[1:])


# ---------------------------------------------------------------------------
# negative coordinates on the separation axes
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("grid", [GridSpec.cartesian("[2:202]*50", "[1:201]*50"),
                                  GridSpec.cartesian("[6:206]*50", "[0:200]*50"),
                                  GridSpec.polar("[3:203]*50", "[0.1:1.1]*20"),
                                  GridSpec.polar("[0:200]*50", "[0.07:1.07]*20"),
                                  GridSpec.observing("[0.0005:0.0205]*20", "[0:0.06]*6")],
                         ids=["cart-2-1", "cart-6-0", "polar-3-0.1", "polar-0-0.07", "obs-0.0005"])
def test_negative_axis_coordinates(dev, ora, zcat, grid):
    run_both(dev, ora, zcat, grid, flags=(0, tos_amd.XI_NO_WINDOW))


@pytest.mark.parametrize("gname", ["cartesian", "polar"])
def test_separation_on_bin_edges(dev, ora, gname):
    """300 sightline pairs, each placed so that its pixel pair at equal
    wavelength (r_par = 0) has r_perp = r within 1e-6 bins of the edge
    4m Mpc/h in the reference's own arithmetic, r_perp^2 = f32(d^2) +
    (d - 2 d c) d (XiEstimator.cc:237-247; the pixel, m and the pair's
    position drawn at random): the reference puts about a quarter of them
    below the edge, and FP32 arithmetic without the re-check bins about a
    quarter differently (a numpy emulation of these pairs).  The FP32 path
    must re-check every such pair and bin it as the reference does
    (Cartesian r_perp axis, Polar r axis)."""
    rng = np.random.default_rng(9400)
    npix, n_pairs, scale = 160, 300, 3000.0
    d = (3400.0 + 0.9 * np.arange(npix)).astype(np.float32)
    lam = (3.56 + 3e-4 * np.arange(npix)).astype(np.float32)
    grid = getattr(GridSpec, gname)()
    max_ang = grid.max_angular_scale(scale)
    ra, dec = [], []
    for _ in range(n_pairs):
        D32 = d[int(rng.integers(npix))]
        D = float(D32)
        edge = 4.0 * int(rng.integers(1, 50))
        c = (float(D32 * D32) + D * D - edge * edge) / (2.0 * D * D)   # f32(D^2) + D^2 - 2 D^2 c = edge^2
        de = float(np.arcsin(rng.uniform(-0.9, 0.9)))
        a0 = float(rng.uniform(0.0, 2.0 * math.pi))
        dra = math.acos((c - math.sin(de) ** 2) / math.cos(de) ** 2)
        assert math.acos(c) < max_ang
        ra += [a0, math.fmod(a0 + dra, 2.0 * math.pi)]
        dec += [de, de]
    n = 2 * n_pairs
    w = rng.uniform(0.5, 2.0, n * npix)
    v = rng.normal(0.0, 0.2, n * npix)
    cat = catalog(ra, dec, [npix] * n, np.tile(d, n), np.tile(lam, n), w, v, scale)
    ref, out = run_both(dev, ora, cat, grid)
    assert out[(tos_amd.FP32C, 0)].num_guard_rechecks >= n_pairs


def test_mu_on_bin_edges(dev, ora):
    """300 sightline pairs, each placed so that one pixel pair has
    mu = r_par / r within 1e-6 bins of the edge m/20 in the reference's
    arithmetic (XiEstimator.cc:183-194; integer distances below 4096, so
    f32(d^2) is exact and either sightline may be the primary): the
    reference puts about a third of them below the edge, FP32 arithmetic
    without the re-check bins about a third differently (a numpy emulation
    of these pairs), and the FP32 path must match the reference."""
    rng = np.random.default_rng(9500)
    npix, n_pairs, scale = 160, 300, 3000.0
    d = (3400.0 + np.arange(npix)).astype(np.float32)
    lam = (3.56 + 3e-4 * np.arange(npix)).astype(np.float32)
    grid = GridSpec.polar()
    max_ang = grid.max_angular_scale(scale)
    ra, dec = [], []
    for _ in range(n_pairs):
        while True:
            k1, k2, m = int(rng.integers(npix)), int(rng.integers(npix)), int(rng.integers(1, 20))
            rpar = abs(float(d[k1]) - float(d[k2]))
            if rpar > 0 and rpar * 20.0 / m < 190.0:
                break
        D1, D2 = float(d[k1]), float(d[k2])
        c = (D1 * D1 + D2 * D2 - (rpar * 20.0 / m) ** 2) / (2.0 * D1 * D2)   # r = rpar / mu
        de = float(np.arcsin(rng.uniform(-0.9, 0.9)))
        a0 = float(rng.uniform(0.0, 2.0 * math.pi))
        dra = math.acos((c - math.sin(de) ** 2) / math.cos(de) ** 2)
        assert math.acos(c) < max_ang
        ra += [a0, math.fmod(a0 + dra, 2.0 * math.pi)]
        dec += [de, de]
    n = 2 * n_pairs
    w = rng.uniform(0.5, 2.0, n * npix)
    v = rng.normal(0.0, 0.2, n * npix)
    cat = catalog(ra, dec, [npix] * n, np.tile(d, n), np.tile(lam, n), w, v, scale)
    ref, out = run_both(dev, ora, cat, grid)
    assert out[(tos_amd.FP32C, 0)].num_guard_rechecks >= n_pairs


@pytest.mark.parametrize("gname", ["cartesian", "polar"])
def test_long_sightlines(dev, ora, gname):
    """Eight sightlines of 3,700-7,000 pixels (a tenth of the usual pixel
    spacing over 900-1500 A rest frame): partner runs longer than a work
    item's partner-pixel cap (4096; such an item holds that one sightline
    pair), windows over thousands of pixels, and transposed tails."""
    cat = synth.lya_catalog(8, area_deg2=0.5, seed=9300, dloglam=3e-5, rest_range=(900.0, 1500.0))
    assert int(np.max(cat.counts())) > 4096
    run_both(dev, ora, cat, getattr(GridSpec, gname)(), flags=(0, tos_amd.XI_NO_WINDOW))


@pytest.mark.parametrize("grid", [GridSpec.polar("[0:50]*100", "[0:1]*20"),
                                  GridSpec.polar("[0:256]*32", "[0:1]*10"),
                                  GridSpec.polar("[0:150]*50", "[0:1]*20")],
                         ids=["polar-w0.5", "polar-w8", "polar-w3"])
def test_polar_r_bin_widths(dev, ora, zcat, grid):
    """r bin widths 0.5, 8 and 3 Mpc/h: guard bands, windows and SIMPLE
    shortcuts derived for power-of-two and other inverse widths."""
    run_both(dev, ora, zcat, grid, flags=(0, tos_amd.XI_NO_WINDOW))


# ---------------------------------------------------------------------------
# Observing theta bin
# ---------------------------------------------------------------------------
def test_observing_theta_axis_above_zero(dev, ora, zcat):
    """theta axis [0.01:0.03]*4 (max_ang 0.03): pairs closer than 0.005 rad
    cast to 2^32 - k, and the unsigned flat index wraps them into the row
    before; pairs in (0.005, 0.01) land in theta bin 0."""
    grid = GridSpec.observing("[0:0.02]*20", "[0.01:0.03]*4")
    ref, _ = run_both(dev, ora, zcat, grid)
    assert ref["stats"]["num_pixel_pairs_used"] > 0


def test_observing_theta_carry(dev, ora, zcat):
    """A pair list built for a larger angle than the theta axis covers
    (tos_pairs_build takes max_ang): theta bins 4 and 5 of a 4-bin axis
    carry into the next dloglam row, and past the last row are skipped,
    exactly as the reference's flat index for such a pair
    (XiEstimator.cc:273-290); the oracle's per-sightline-pair function on
    the same pair list is the reference."""
    grid = GridSpec.observing("[0:0.02]*20", "[0:0.02]*4")
    max_ang = 0.03
    dcat = dev.upload(zcat)
    pairs = dcat.pairs(max_ang, order=4, nthreads=8)
    row_los, _, row_ptr, other, cs = pairs.download()
    nb = grid.nbins_total
    ref = {"num_pairs": np.zeros(nb, np.int64), "didj": np.zeros(nb), "wgt": np.zeros(nb)}
    used = 0
    for r, i in enumerate(row_los):
        for q in range(row_ptr[r], row_ptr[r + 1]):
            xi, u = ora.accumulate_los_pair(zcat, grid, int(i), int(other[q]), float(cs[q]))
            used += u
            for f in ("num_pairs", "didj", "wgt"):
                ref[f] += xi[f]
    sep = np.arccos(cs)
    assert (sep >= 0.02).any() and (sep < 0.02).any()
    for prec in (tos_amd.FP64, tos_amd.FP32C):
        for f in (0, tos_amd.XI_NO_ZCONST):
            print(f"case: theta carry precision {prec} flags {f}", flush=True)
            h = pairs.run(grid, precision=prec, flags=f)
            check(h, ref)
            assert h.num_pixel_pairs_used == used
    pairs.free()
    dcat.free()
