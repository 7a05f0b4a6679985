"""Pins of the CPU oracle against facts fixed outside it: x86 conversion
semantics, closed-form geometry computed from 3D vectors in extended
precision, hand-counted lattice fixtures from the reference's sample-data,
HEALPix round trips and invariants of the estimator.  The reference ships no
golden xi values (it has no tests), so these checks are what fix the oracle.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

LD = np.longdouble


# ---------------------------------------------------------------------------
# x86-64 static_cast<unsigned>(double)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("x,expect", [
    (float("nan"), 0), (-0.5, 0), (-0.0, 0), (0.99, 0), (5.7, 5), (-1.0, 4294967295),
    (-1.5, 4294967295), (-4.9, 4294967292), (2.0 ** 32 + 5, 5), (1e19, 0), (-1e19, 0),
    (-(2.0 ** 31), 2 ** 31), (4294967295.5, 4294967295)])
def test_cvt_u32_matches_x86(ora, x, expect):
    assert ora.cvt_u32(x) == expect


# ---------------------------------------------------------------------------
# XiBin (types.h:69-121)
# ---------------------------------------------------------------------------
def test_accumulate_and_finalize_closed_form(ora):
    b = ora.XiBin()
    L = ora.lib()
    L.ora_accumulate_pair(b, 0.5, 3.5, 2.0, -0.25, 3.75, 4.0)   # v1 l1 w1 v2 l2 w2
    assert (b.didj, b.di, b.dj, b.wgt, b.wi, b.wj, b.z, b.num_pairs) == (-1.0, 1.0, -1.0, 8.0, 2.0, 4.0, 3.625, 1)
    L.ora_accumulate_pair(b, 0.1, 3.6, 3.0, 0.3, 3.61, 7.0)
    f32 = np.float32
    wdi, wdj = f32(3.0) * f32(0.1), f32(7.0) * f32(0.3)
    assert b.didj == -1.0 + float(f32(wdi * wdj))
    assert b.wgt == 8.0 + float(f32(21.0))
    assert b.z == 3.625 + 0.5 * float(f32(f32(3.6) + f32(3.61)))
    L.ora_finalize(b)
    assert b.didj == pytest.approx((-1.0 + float(f32(wdi * wdj))) / 29.0, rel=1e-15)
    assert b.di == pytest.approx((1.0 + float(wdi)) / 5.0, rel=1e-15)
    assert b.z == pytest.approx((3.625 + 0.5 * float(f32(f32(3.6) + f32(3.61)))) / 2, rel=1e-15)
    e = ora.XiBin()
    L.ora_finalize(e)
    assert (e.didj, e.di, e.dj, e.z) == (0.0, 0.0, 0.0, 0.0)


# ---------------------------------------------------------------------------
# angularSeparation (types.h:37-39) vs a 3D dot product
# ---------------------------------------------------------------------------
def unit(ra, dec):
    ra, dec = LD(ra), LD(dec)
    return np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], dtype=LD)


def test_cos_separation_is_the_dot_product(ora):
    rng = np.random.default_rng(3)
    for _ in range(500):
        ra1, ra2 = rng.uniform(0, 2 * np.pi, 2)
        dec1, dec2 = rng.uniform(-1.5, 1.5, 2)
        c = ora.cos_separation(ra1, dec1, ra2, dec2)
        assert abs(c - float(unit(ra1, dec1) @ unit(ra2, dec2))) < 4e-16
    assert ora.cos_separation(0.0, 0.0, 0.0, 0.0) == 1.0
    # RA wraps: 0.01 and 2pi - 0.01 are 0.02 apart on the equator
    assert ora.cos_separation(0.01, 0.0, 2 * np.pi - 0.01, 0.0) == pytest.approx(math.cos(0.02), abs=1e-15)


# ---------------------------------------------------------------------------
# HEALPix RING
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64])
def test_healpix_round_trip(ora, nside):
    npix = 12 * nside * nside
    for p in range(npix):
        th, ph = ora.pix2ang_ring(nside, p)
        assert ora.ang2pix_ring(nside, th, ph) == p
    # pixel 0 touches the north pole, the last one the south pole
    assert ora.ang2pix_ring(nside, 1e-9, 0.1) == 0
    assert ora.ang2pix_ring(nside, np.pi - 1e-9, 0.1) == npix - 4
    # first equatorial-belt ring starts at 2 nside (nside - 1)
    th, _ = ora.pix2ang_ring(nside, 2 * nside * (nside - 1))
    assert abs(np.cos(th) - 2.0 / 3.0) < 1e-12 if nside > 1 else True


def test_healpix_equal_area(ora):
    rng = np.random.default_rng(9)
    n, nside = 200000, 4
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    pix = np.array([ora.ang2pix_ring(nside, math.acos(a), b) for a, b in zip(z, phi)])
    counts = np.bincount(pix, minlength=192)
    mean = n / 192
    assert np.all(np.abs(counts - mean) < 6 * math.sqrt(mean))


def test_product_healpix_matches_oracle(ora):
    from tos_amd import healpix_radec2pix
    rng = np.random.default_rng(4)
    ra = rng.uniform(0, 2 * np.pi, 20000)
    dec = np.arcsin(rng.uniform(-1, 1, 20000))
    for order in (0, 2, 4, 6):
        got = healpix_radec2pix(order, ra, dec)
        want = np.array([ora.lib().ora_radec2pix(1 << order, a, b) for a, b in zip(ra, dec)])
        assert np.array_equal(got, want)


# ---------------------------------------------------------------------------
# 3D brute force on the reference's lattice fixtures
# ---------------------------------------------------------------------------
def load_dat(name):
    a = np.loadtxt(os.path.join(GOLDEN, name))
    return [a[:, k].copy() for k in range(5)]


def test_brute_nanoer_hand_counts(ora):
    x, y, z, d, w = load_dat("delta-nanoer.dat")   # 2x2x2 lattice, spacing 5
    ds, ws, cn = ora.brute3d_f64(x, y, z, d, w, 0.0, 100.0, 50)
    expect = np.zeros(50, np.int64)
    expect[2] = 12   # edges, r = 5
    expect[3] = 12   # face diagonals, r = 7.07
    expect[4] = 4    # body diagonals, r = 8.66
    assert np.array_equal(cn, expect)
    # sums by distance class, computed directly
    pts = np.stack([x, y, z], 1)
    for b, r2 in ((2, 25.0), (3, 50.0), (4, 75.0)):
        s = sum(d[i] * d[j] for i in range(8) for j in range(i + 1, 8)
                if abs(((pts[i] - pts[j]) ** 2).sum() - r2) < 1e-9)
        assert ds[b] == pytest.approx(s, rel=1e-14)
        assert ws[b] == expect[b]
    # brute.cu float binning: slot = bin + 1, slot 0 / nbins+1 under/overflow
    f = [a.astype(np.float32) for a in (x, y, z, d, w)]
    ds32, ws32, cn32 = ora.brute3d_f32(*f, 0.0, 100.0, 50)
    assert np.array_equal(cn32[1:51], expect) and cn32[0] == 0 and cn32[51] == 0


def test_brute_nano_vs_lattice_distances(ora):
    x, y, z, d, w = load_dat("delta-nano.dat")     # 4x4x4 lattice
    n = len(x)
    ds, ws, cn = ora.brute3d_f64(x, y, z, d, w, 0.0, 100.0, 50)
    # lattice separations are 5 sqrt(k): bins from exact integer arithmetic
    ijk = np.round((np.stack([x, y, z], 1) - 2.5) / 5).astype(int)
    expect = np.zeros(50, np.int64)
    dsum = np.zeros(50)
    for i in range(n):
        for j in range(i + 1, n):
            k = int(((ijk[i] - ijk[j]) ** 2).sum())
            # bin = floor(5 sqrt(k) / 2) = largest m with (2m)^2 <= 25 k
            m = math.isqrt(25 * k) // 2
            if (2 * m + 2) ** 2 <= 25 * k:
                m += 1
            expect[m] += 1
            dsum[m] += d[i] * d[j]
    assert np.array_equal(cn, expect)
    np.testing.assert_allclose(ds, dsum, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# Sightline-pair geometry (XiEstimator.cc:161-296) from 3D vectors
# ---------------------------------------------------------------------------
class Cat:
    def __init__(self, ra, dec, counts, d, l, w, v):
        self.ra = np.asarray(ra, float)
        self.dec = np.asarray(dec, float)
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.distance = np.asarray(d, np.float32)
        self.loglam = np.asarray(l, np.float32)
        self.weight = np.asarray(w, np.float32)
        self.value = np.asarray(v, np.float32)


def two_los(sep, seed=0, n=140):
    rng = np.random.default_rng(seed)
    ra = [0.3, 0.3 + sep / math.cos(0.1)]
    dec = [0.1, 0.1 + 0.3 * sep]
    kk = np.sort(rng.choice(600, size=2 * n, replace=False)).reshape(2, n)
    d = np.concatenate([3600.0 + 1.91 * kk[0], 3580.0 + 1.93 * kk[1]])
    l = np.concatenate([3.5563 + 3e-4 * kk[0], 3.5563 + 3e-4 * kk[1]])
    w = rng.uniform(3.4, 23.5, 2 * n)
    v = rng.normal(0, 0.2, 2 * n)
    return Cat(ra, dec, [n, n], d, l, w, v)


def geometry(cat):
    """Per pixel pair: r_par, r_perp^2 and r^2 from 3D positions in long
    double, plus e1 = f32(d1^2) - d1^2 (the reference squares d1 in float,
    XiEstimator.cc:237)."""
    u = [unit(cat.ra[k], cat.dec[k]) for k in range(2)]
    a0, a1, b1 = cat.offsets[0], cat.offsets[1], cat.offsets[2]
    d1 = cat.distance[a0:a1].astype(LD)
    d2 = cat.distance[a1:b1].astype(LD)
    p1 = d1[:, None, None] * u[0][None, None, :]
    p2 = d2[None, :, None] * u[1][None, None, :]
    r2 = ((p1 - p2) ** 2).sum(-1)
    e1 = (cat.distance[a0:a1] * cat.distance[a0:a1]).astype(LD) - d1 * d1
    r2 = r2 + e1[:, None]
    rpar = np.abs(d1[:, None] - d2[None, :])
    return rpar, r2 - rpar ** 2, r2


def terms(cat):
    a0, a1, b1 = cat.offsets[0], cat.offsets[1], cat.offsets[2]
    wd1 = cat.weight[a0:a1] * cat.value[a0:a1]
    wd2 = cat.weight[a1:b1] * cat.value[a1:b1]
    didj = (wd1[:, None] * wd2[None, :]).astype(np.float64)
    ww = (cat.weight[a0:a1][:, None] * cat.weight[a1:b1][None, :]).astype(np.float64)
    return didj, ww


def compare(xi, used, idx, ok, amb, cat, nb):
    """Unambiguous pairs must land exactly where idx says; ambiguous pairs
    (within 1e-6 of an edge) may land in either neighbour."""
    didj, ww = (t.ravel() for t in terms(cat))
    cnt = np.bincount(idx[ok & ~amb], minlength=nb)
    slack = np.zeros(nb, np.int64)
    for b in idx[amb & ok]:
        slack[max(b - 1, 0):b + 2] += 1
    got = xi["num_pairs"]
    assert np.all(got >= cnt - slack) and np.all(got <= cnt + slack)
    assert amb.sum() < 0.005 * amb.size
    clean = slack == 0
    sd = np.bincount(idx[ok & ~amb], weights=didj[ok & ~amb], minlength=nb)
    sw = np.bincount(idx[ok & ~amb], weights=ww[ok & ~amb], minlength=nb)
    np.testing.assert_allclose(xi["didj"][clean], sd[clean], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(xi["wgt"][clean], sw[clean], rtol=1e-12, atol=1e-9)
    assert used == int(got.sum())


def near(t, eps=1e-6):
    return np.abs(t - np.round(t)) < eps


@pytest.mark.parametrize("sep", [0.003, 0.02, 0.05])
def test_cartesian_geometry(ora, sep):
    from tos_amd import GridSpec
    cat = two_los(sep, seed=int(sep * 1000))
    grid = GridSpec.cartesian()
    c = ora.cos_separation(cat.ra[0], cat.dec[0], cat.ra[1], cat.dec[1])
    xi, used = ora.accumulate_los_pair(cat, grid, 0, 1, c)
    rpar, rperp2, _ = geometry(cat)
    tpar = rpar / 4
    tperp = np.sqrt(np.abs(rperp2)) / 4
    ok = (tpar < 50) & (tperp < 50)
    idx = (np.floor(tpar).astype(np.int64) * 50 + np.floor(np.minimum(tperp, 60)).astype(np.int64)).ravel()
    # r_par = |f32(d1 - d2)| is exact in both computations; only r_perp can
    # sit within rounding of an edge
    amb = near(tperp).ravel()
    compare(xi, used, idx, ok.ravel(), amb, cat, 2500)


@pytest.mark.parametrize("sep", [0.003, 0.02, 0.05])
def test_polar_geometry(ora, sep):
    from tos_amd import GridSpec
    cat = two_los(sep, seed=7 + int(sep * 1000))
    grid = GridSpec.polar()
    c = ora.cos_separation(cat.ra[0], cat.dec[0], cat.ra[1], cat.dec[1])
    xi, used = ora.accumulate_los_pair(cat, grid, 0, 1, c)
    rpar, _, r2 = geometry(cat)
    r = np.sqrt(r2)
    tr = r / 4
    tmu = rpar / r * 20
    ok = (tr < 50) & (tmu < 20)
    idx = (np.floor(np.minimum(tr, 60)).astype(np.int64) * 20 +
           np.floor(np.minimum(tmu, 30)).astype(np.int64)).ravel()
    amb = (near(tr) | near(tmu)).ravel()
    compare(xi, used, idx, ok.ravel(), amb, cat, 1000)


def test_observing_geometry(ora):
    from tos_amd import GridSpec
    cat = two_los(0.02, seed=3)
    grid = GridSpec.observing("[0:0.02]*20", "[0:0.06]*6")
    c = ora.cos_separation(cat.ra[0], cat.dec[0], cat.ra[1], cat.dec[1])
    xi, used = ora.accumulate_los_pair(cat, grid, 0, 1, c)
    u0, u1 = unit(cat.ra[0], cat.dec[0]), unit(cat.ra[1], cat.dec[1])
    theta = float(np.arctan2(np.linalg.norm(np.cross(u0.astype(float), u1.astype(float))), float(u0 @ u1)))
    ith = int(theta / 0.01)
    l1 = cat.loglam[:140].astype(LD)
    l2 = cat.loglam[140:].astype(LD)
    tdl = np.abs(l1[:, None] - l2[None, :]) / LD(0.001)
    ok = tdl < 20
    idx = (np.floor(np.minimum(tdl, 30)).astype(np.int64) * 6 + ith).ravel()
    compare(xi, used, idx, ok.ravel(), near(tdl).ravel(), cat, 120)


# ---------------------------------------------------------------------------
# Reference quirks reproduced on purpose
# ---------------------------------------------------------------------------
def float_with_e1(sign):
    """A float distance whose float square is below (sign<0) / above (sign>0)
    the exact square."""
    d = np.float32(4000.0)
    while True:
        e1 = float(np.float64(d * d)) - float(d) ** 2
        if (sign < 0 and e1 < 0) or (sign > 0 and e1 > 0) or (sign == 0 and e1 == 0):
            return d
        d = np.nextafter(d, np.float32(5000.0))


@pytest.mark.parametrize("sign,expect_used", [(-1, 1), (1, 1), (0, 0)])
def test_polar_nan_path(ora, sign, expect_used):
    """r_sq = f32(d^2) - d^2 < 0 for coincident sightlines: sqrt gives NaN,
    the cast gives bin 0, 'r == 0' is false, mu is NaN -> bin (0, 0)
    (XiEstimator.cc:183-192).  r_sq == 0 is skipped by 'r == 0'."""
    from tos_amd import GridSpec
    d = float(float_with_e1(sign)) if sign != 0 else 4096.0
    cat = Cat([0.0, 0.0], [0.0, 0.0], [1, 1], [d, d], [3.6, 3.6], [1.0, 2.0], [0.5, 0.5])
    c = ora.cos_separation(0.0, 0.0, 0.0, 0.0)
    assert c == 1.0
    xi, used = ora.accumulate_los_pair(cat, GridSpec.polar(), 0, 1, c)
    assert used == expect_used
    assert int(xi["num_pairs"][0]) == expect_used


def test_z_axis_quirks(ora):
    """The z axis is cast without a range check and scaled by 1/NBins
    (XiEstimator.cc:232,251-253): a z axis just above the data still maps
    to bin 0; one more than a unit above wraps to 2^32-1, which shifts
    every flat index down by one; with 2 z bins every pair lands in z bin 0."""
    from tos_amd import GridSpec, synth
    cat = synth.lya_catalog(60, area_deg2=6.0, seed=21)
    base = ora.xi_run(cat, GridSpec.cartesian(), cat.scale, order=4, nthreads=4)["raw"]["num_pairs"]
    above = ora.xi_run(cat, GridSpec.cartesian(z="[3.8:4.0]*1"), cat.scale, order=4, nthreads=4)["raw"]["num_pairs"]
    assert np.array_equal(above, base)
    wrap = ora.xi_run(cat, GridSpec.cartesian(z="[5:6]*1"), cat.scale, order=4, nthreads=4)["raw"]["num_pairs"]
    assert np.array_equal(wrap[:-1], base[1:]) and wrap[-1] == 0
    two = ora.xi_run(cat, GridSpec.cartesian(z="[3.556:3.735]*2"), cat.scale, order=4, nthreads=4)["raw"]["num_pairs"]
    assert np.array_equal(two.reshape(2500, 2)[:, 0], base) and not two.reshape(2500, 2)[:, 1].any()


# ---------------------------------------------------------------------------
# Estimator invariants
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def small_cat():
    from tos_amd import synth
    return synth.lya_catalog(150, area_deg2=15.0, seed=77)


def test_counts_sum_to_pixel_pairs_used(ora, small_cat):
    from tos_amd import GridSpec
    r = ora.xi_run(small_cat, GridSpec.cartesian(), small_cat.scale, order=4, nthreads=4)
    assert int(r["raw"]["num_pairs"].sum()) == r["stats"]["num_pixel_pairs_used"]
    assert r["stats"]["num_pixels"] == small_cat.n_pix
    assert r["stats"]["num_pixel_pairs"] >= r["stats"]["num_pixel_pairs_used"]


def test_sky_binning_independence(ora, small_cat):
    """A superset neighbour search plus the exact cos cut makes the totals
    independent of the HEALPix order (only subsamples depend on it)."""
    from tos_amd import GridSpec
    g = GridSpec.polar()
    runs = [ora.xi_run(small_cat, g, small_cat.scale, order=o, nthreads=4) for o in (1, 3, 4, 6)]
    for r in runs[1:]:
        assert np.array_equal(r["raw"]["num_pairs"], runs[0]["raw"]["num_pairs"])
        np.testing.assert_allclose(r["raw"]["didj"], runs[0]["raw"]["didj"], rtol=1e-11, atol=1e-9)
        assert r["stats"]["num_pixel_pairs"] == runs[0]["stats"]["num_pixel_pairs"]
    # subsamples add up to the total
    r = ora.xi_run(small_cat, g, small_cat.scale, order=4, nthreads=4, per_skybin=True)
    assert np.array_equal(r["per_skybin"]["num_pairs"].sum(0), r["raw"]["num_pairs"])


def test_observing_grid_relabel_invariance(ora, small_cat):
    """The Observing branch is symmetric in (p1, p2), so relabelling the
    sightlines (which swaps primary and partner) leaves counts unchanged."""
    from tos_amd import GridSpec
    g = GridSpec.observing()
    perm = np.random.default_rng(1).permutation(small_cat.n_los)
    a = ora.xi_run(small_cat, g, small_cat.scale, order=4, nthreads=4)["raw"]
    b = ora.xi_run(small_cat.subset(perm), g, small_cat.scale, order=4, nthreads=4)["raw"]
    assert np.array_equal(a["num_pairs"], b["num_pairs"])
    np.testing.assert_allclose(a["wgt"], b["wgt"], rtol=1e-12)


def test_plate_bins_annulus(ora):
    """PlateBins neighbours: plates whose centre lies within `radius` of the
    1.5 deg annulus around the primary (PlateBins.h:74-85).  Hand case: three
    sightlines on plate 1 (centre at the origin), one on plate 2 whose centre
    is 2 deg away; radius 0.6 deg."""
    from tos_amd import GridSpec
    deg = math.pi / 180
    ra = np.array([0.0, 0.2, 0.4, 1.9]) * deg
    dec = np.zeros(4)
    counts = [3, 3, 3, 3]
    n = 12
    cat = Cat(ra, dec, counts, np.linspace(3700, 3720, n), np.full(n, 3.6), np.ones(n), np.ones(n))
    plate_of = np.array([1, 1, 1, 2])
    plates = (plate_of, np.array([2, 1]), np.array([2.0 * deg, 0.0]), np.array([0.0, 0.0]))
    grid = GridSpec.cartesian()
    # the cut uses max_ang = 200 / scale; pick scale so max_ang = 0.6 deg
    scale = 200.0 / (0.6 * deg)
    r = ora.xi_run(cat, grid, scale, plates=plates, nthreads=1)
    # primary 0 (ra 0): plate 1 centre at 0 deg -> |0 - 1.5| = 1.5 > 0.6: not a
    # neighbour; plate 2 centre at 2 deg -> |2 - 1.5| = 0.5 < 0.6: neighbour, but
    # LOS 3 is 1.9 deg away > 0.6 -> cut.  Primaries 1 and 2 likewise see only
    # plate 2, whose member is beyond the cut.  So the pairs (0,1) (0,2) (1,2),
    # all closer than 0.6 deg, are never considered: nothing is counted.
    assert r["stats"]["num_sightline_pairs_used"] == 0
    # with a 1.6 deg radius every plate is a neighbour and every pair within 1.6 deg counts
    scale = 200.0 / (1.6 * deg)
    r = ora.xi_run(cat, grid, scale, plates=plates, nthreads=1)
    # pairs within 1.6 deg: (0,1) (0,2) (1,2) (2,3); (1,3) is 1.7 deg, (0,3) 1.9 deg
    assert r["stats"]["num_sightline_pairs_used"] == 4
