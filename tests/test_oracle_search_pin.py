"""Pin of the oracle's candidate search (XiEstimator.cc:121-159) by brute
force: every sightline pair i < j whose cos separation, in the operation
order of types.h:37-39, exceeds cos(max_ang) (XiEstimator.cc:147: strictly
greater), attributed to the sky bin of the primary i.  The brute force
enumerates all N(N-1)/2 pairs with numpy and owes nothing to the oracle's
declination window or to HEALPix cells, so a wrong `j <= i`, a `>=` cut or
a window that is too narrow fails here.  Catalogues carry partners placed
at cos(max_ang) - 1 ulp, exactly at it and + 1 ulp, on the edge of the
oracle's declination window."""
import math

import numpy as np
import pytest


def trig(cat):
    # host libm (math.sin/cos call the C library, as the reference's
    # SkyObject constructor does, types.h:20-26)
    sd = np.array([math.sin(d) for d in cat.dec])
    cd = np.array([math.cos(d) for d in cat.dec])
    sr = np.array([math.sin(r) for r in cat.ra])
    cr = np.array([math.cos(r) for r in cat.ra])
    return sd, cd, sr, cr


def cos_sep_all(cat):
    """c[i, j] = sd_i sd_j + (cd_i cd_j) (sr_i sr_j + cr_i cr_j), each
    product and sum rounded once (numpy evaluates into temporaries: no
    contraction)."""
    sd, cd, sr, cr = trig(cat)
    t1 = sd[:, None] * sd[None, :]
    t2 = cd[:, None] * cd[None, :]
    t3 = sr[:, None] * sr[None, :]
    t4 = cr[:, None] * cr[None, :]
    return t1 + t2 * (t3 + t4)


def cos_sep_one(ra1, dec1, ra2, dec2):
    t1 = math.sin(dec1) * math.sin(dec2)
    t2 = math.cos(dec1) * math.cos(dec2)
    t3 = math.sin(ra1) * math.sin(ra2)
    t4 = math.cos(ra1) * math.cos(ra2)
    return t1 + t2 * (t3 + t4)


def near_cut_offsets(ra0, dec0, cos_max, max_ang):
    """Declination offsets d (same right ascension) whose cos separation is
    cos_max - 1 ulp, cos_max and cos_max + 1 ulp, found by walking d over
    consecutive doubles around max_ang."""
    want = {math.nextafter(cos_max, -1.0): None, cos_max: None, math.nextafter(cos_max, 2.0): None}
    step = math.ulp(max_ang) * 8
    for k in range(-20000, 20000):
        dd = max_ang + k * step
        c = cos_sep_one(ra0, dec0, ra0, dec0 + dd)
        if c in want and want[c] is None:
            want[c] = dd
    return [v for v in want.values() if v is not None], want


def brute(ora, cat, grid, max_ang, bin_of_los):
    cos_max = math.cos(max_ang)
    c = cos_sep_all(cat)
    n = cat.n_los
    iu, ju = np.triu_indices(n, k=1)
    keep = c[iu, ju] > cos_max
    pi, pj = iu[keep], ju[keep]
    cnt = np.diff(cat.offsets)
    stats = {"num_sightline_pairs_used": int(keep.sum()),
             "num_pixel_pairs": int((cnt[pi] * cnt[pj]).sum())}
    keys = np.unique(bin_of_los)
    nb = grid.nbins_total
    per = {int(k): np.zeros(nb, dtype=ora.XIBIN_DTYPE) for k in keys}
    used = 0
    for i, j in zip(pi, pj):
        xi, u = ora.accumulate_los_pair(cat, grid, int(i), int(j), float(c[i, j]))
        used += u
        tgt = per[int(bin_of_los[i])]
        for f in ("didj", "wgt", "num_pairs"):
            tgt[f] += xi[f]
    stats["num_pixel_pairs_used"] = used
    return stats, per


def near_cut_catalog(order_seed=0):
    """A small seeded catalogue plus three partners of sightline 0 at the
    cut, and a duplicate of sightline 1 at its exact position (c = 1)."""
    from tos_amd import GridSpec, synth
    base = synth.lya_catalog(160, area_deg2=16.0, seed=910 + order_seed)
    grid = GridSpec.cartesian()
    max_ang = grid.max_angular_scale(base.scale)
    cos_max = math.cos(max_ang)
    offs, found = near_cut_offsets(float(base.ra[0]), float(base.dec[0]), cos_max, max_ang)
    extra = [0] * len(offs) + [1]
    cat = base.subset(np.concatenate([np.arange(base.n_los), np.array(extra, dtype=np.int64)]))
    n0 = base.n_los
    for k, d in enumerate(offs):
        cat.ra[n0 + k] = base.ra[0]
        cat.dec[n0 + k] = base.dec[0] + d
    cat.scale = base.scale
    return cat, grid, max_ang, found


def test_near_cut_offsets_exist():
    """The construction reaches all three cos values around the cut."""
    from tos_amd import GridSpec, synth
    base = synth.lya_catalog(10, area_deg2=16.0, seed=910)
    max_ang = GridSpec.cartesian().max_angular_scale(base.scale)
    offs, found = near_cut_offsets(float(base.ra[0]), float(base.dec[0]), math.cos(max_ang), max_ang)
    assert len(offs) == 3, found


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5, 6])
def test_oracle_search_equals_brute_force(ora, order):
    cat, grid, max_ang, found = near_cut_catalog(order)
    assert ora.max_angular_scale(grid, cat.scale) == max_ang
    bol = ora.healpix_bins(cat, order)
    r = ora.xi_run(cat, grid, cat.scale, order=order, nthreads=4, per_skybin=True)
    st, per = brute(ora, cat, grid, max_ang, bol)
    for k in ("num_sightline_pairs_used", "num_pixel_pairs", "num_pixel_pairs_used"):
        assert r["stats"][k] == st[k], k
    keys = [int(k) for k in r["skybin_keys"]]
    assert keys == sorted(per)
    for k, h in zip(keys, r["per_skybin"]):
        assert np.array_equal(h["num_pairs"], per[k]["num_pairs"]), k
        np.testing.assert_allclose(h["didj"], per[k]["didj"], rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(h["wgt"], per[k]["wgt"], rtol=1e-12, atol=1e-9)


def test_brute_force_sees_the_cut_cases(ora):
    """The partner exactly at cos_max is rejected and the one 1 ulp above is
    kept (strict '>', XiEstimator.cc:147); the duplicate at c = 1 is kept
    once, as the pair (1, copy), never as a self pair (1, 1)."""
    cat, grid, max_ang, found = near_cut_catalog(0)
    cos_max = math.cos(max_ang)
    c = cos_sep_all(cat)
    n0 = cat.n_los - 4
    assert c[0, n0 + 0] == math.nextafter(cos_max, -1.0)
    assert c[0, n0 + 1] == cos_max
    assert c[0, n0 + 2] == math.nextafter(cos_max, 2.0)
    # a sightline and its copy: the reference formula gives 1 - 1 ulp here,
    # far above the cut (a self pair would look the same)
    assert c[1, n0 + 3] > math.nextafter(cos_max, 2.0) and c[1, 1] > cos_max
    st, _ = brute(ora, cat, grid, max_ang, np.zeros(cat.n_los, dtype=np.int32))
    base = ora.xi_run(cat.subset(np.arange(n0)), grid, cat.scale, order=4, nthreads=4)["stats"]
    r = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=4)["stats"]
    assert r["num_sightline_pairs_used"] == st["num_sightline_pairs_used"]
    assert r["num_sightline_pairs_used"] > base["num_sightline_pairs_used"]
    # the pairs of sightline 0 with the three partners at the cut: only the
    # one 1 ulp above cos_max counts
    iu, ju = np.triu_indices(cat.n_los, k=1)
    keep = c[iu, ju] > cos_max
    kept0 = set(ju[keep & (iu == 0)].tolist())
    assert n0 + 2 in kept0 and n0 + 0 not in kept0 and n0 + 1 not in kept0
