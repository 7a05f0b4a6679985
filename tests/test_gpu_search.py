"""Candidate search on the GPU (tos_pairs_build with on_device = 1, HEALPix
sky bins; the skybin_xi_task loop of XiEstimator.cc:121-159): rows,
partners, cos values and statistics equal the host search bit for bit, the
statistics equal the oracle's, and the histograms built on the device pair
list match the oracle."""
import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth

pytestmark = pytest.mark.gpu


def same_pairs(a, b):
    for x, y, name in zip(a.download(), b.download(), ("row_los", "row_skybin", "row_ptr", "other", "cos_sep")):
        assert x.dtype == y.dtype and x.shape == y.shape, name
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name
    assert a.stats() == b.stats()


@pytest.fixture(scope="module")
def mini():
    return synth.lya_catalog(330, area_deg2=33.0, seed=2001)


@pytest.mark.parametrize("order", [0, 2, 4, 6])
@pytest.mark.parametrize("ang_scale", [0.25, 1.0, 3.0])
def test_device_search_equals_host(dev, mini, order, ang_scale):
    # order 0 puts every sightline in one or two bins (rows of several
    # 64-member chunks); order 6 spreads them over many small bins
    grid = GridSpec.cartesian()
    max_ang = ang_scale * grid.max_angular_scale(mini.scale)
    print(f"order={order} max_ang={max_ang:.6g}", flush=True)
    dcat = dev.upload(mini)
    host = dcat.pairs(max_ang, order=order, on_device=False)
    devp = dcat.pairs(max_ang, order=order, on_device=True)
    same_pairs(devp, host)
    devp.free()
    host.free()
    dcat.free()


def test_device_search_stats_and_xi_match_oracle(dev, ora, mini):
    grid = GridSpec.cartesian()
    ref = ora.xi_run(mini, grid, mini.scale, order=4, nthreads=4)
    dcat = dev.upload(mini)
    pairs = dcat.pairs(grid.max_angular_scale(mini.scale), order=4, on_device=True)
    st = pairs.stats()
    for k, v in ref["stats"].items():
        if k in st:
            assert st[k] == v, k
    for prec in (tos_amd.FP64, tos_amd.FP32C):
        h = pairs.run(grid, precision=prec)
        assert np.array_equal(h.num_pairs.astype(np.int64), ref["raw"]["num_pairs"])
        m = ref["raw"]["wgt"] > 0
        assert np.allclose(h.wgt[m], ref["raw"]["wgt"][m], rtol=1e-10, atol=0)
        assert np.allclose(h.didj[m], ref["raw"]["didj"][m], rtol=1e-10, atol=1e-300)
    pairs.free()
    dcat.free()


def test_device_search_empty_and_tiny(dev):
    # one sightline: no candidate pairs, one occupied bin
    one = synth.lya_catalog(1, area_deg2=1.0, seed=7)
    dcat = dev.upload(one)
    p = dcat.pairs(0.01, order=4, on_device=True)
    st = p.stats()
    assert st["n_rows"] == 0 and st["num_sightline_pairs_used"] == 0 and st["n_skybins"] == 1
    h = p.run(GridSpec.cartesian())
    assert int(h.num_pairs.sum()) == 0
    p.free()
    dcat.free()
    # max_ang = 0: every candidate is rejected (cos(sep) > 1 never holds)
    two = synth.lya_catalog(20, area_deg2=1.0, seed=8)
    dcat = dev.upload(two)
    a = dcat.pairs(0.0, order=3, on_device=True)
    b = dcat.pairs(0.0, order=3, on_device=False)
    same_pairs(a, b)
    assert a.stats()["num_sightline_pairs_used"] == 0
    a.free()
    b.free()
    dcat.free()


@pytest.mark.parametrize("spacing,axis", [(2.5, "[0:200]*50"), (1.0, "[0:40]*10"), (0.7, "[0:120]*30")])
def test_plate_search_device_equals_host_and_oracle(dev, ora, mini, spacing, axis):
    """PlateBins (PlateBins.h:74-85): the GPU plate search against the host
    search bit for bit, and the histograms against the oracle.  The smaller
    grids give max_ang < 1.5 deg, where the annulus test excludes the plates
    closest to a sightline (the reference's behaviour, kept)."""
    plates = synth.plate_tiling(mini, spacing_deg=spacing)
    grid = GridSpec.cartesian(axis, axis)
    max_ang = grid.max_angular_scale(mini.scale)
    print(f"plates spacing={spacing} axis={axis} max_ang={max_ang:.5f}", flush=True)
    dcat = dev.upload(mini)
    host = dcat.pairs(max_ang, plates=plates, on_device=False)
    devp = dcat.pairs(max_ang, plates=plates, on_device=True)
    same_pairs(devp, host)
    ref = ora.xi_run(mini, grid, mini.scale, plates=plates, nthreads=4)
    st = devp.stats()
    assert st["num_sightline_pairs_used"] == ref["stats"]["num_sightline_pairs_used"]
    assert st["num_pixel_pairs"] == ref["stats"]["num_pixel_pairs"]
    assert st["num_sightline_pairs"] == ref["stats"]["num_sightline_pairs_considered"]
    for prec in (tos_amd.FP64, tos_amd.FP32C):
        h = devp.run(grid, precision=prec)
        assert np.array_equal(h.num_pairs.astype(np.int64), ref["raw"]["num_pairs"])
        m = ref["raw"]["wgt"] > 0
        assert np.allclose(h.wgt[m], ref["raw"]["wgt"][m], rtol=1e-10, atol=0)
    devp.free()
    host.free()
    dcat.free()


@pytest.mark.parametrize("order", [0, 3, 4, 6])
def test_product_searches_equal_brute_force_at_the_cut(dev, ora, order):
    """The product's candidate searches (host and device) on catalogues with
    partners at cos(max_ang) - 1 ulp, exactly at it and + 1 ulp, and a copy
    of a sightline (c ~ 1): the pair statistics and per-sky-bin histograms
    of an O(N^2) enumeration of every i < j with c > cos_max, so neither
    neighbour-cell search drops a pair the reference keeps."""
    from test_oracle_search_pin import brute, near_cut_catalog
    cat, grid, max_ang, _ = near_cut_catalog(order)
    st, per = brute(ora, cat, grid, max_ang, ora.healpix_bins(cat, order))
    dcat = dev.upload(cat)
    for on_device in (False, True):
        print(f"case: near-cut catalogue, order {order}, device search {on_device}", flush=True)
        pairs = dcat.pairs(max_ang, order=order, on_device=on_device)
        s = pairs.stats()
        assert s["num_sightline_pairs_used"] == st["num_sightline_pairs_used"]
        assert s["num_pixel_pairs"] == st["num_pixel_pairs"]
        tot, keys, sn, sd, sw = pairs.run_per_skybin(grid, precision=tos_amd.FP64)
        assert tot.num_pixel_pairs_used == st["num_pixel_pairs_used"]
        for k, h in per.items():
            i = int(np.searchsorted(keys, k))
            if i < len(keys) and keys[i] == k:
                assert np.array_equal(sn[i].astype(np.int64), h["num_pairs"]), k
            else:
                assert not h["num_pairs"].any(), k
        pairs.free()
    dcat.free()
