"""Covariance of the sky-bin subsamples on the GPU (tos_subsample_covariance,
XiEstimator.cc:60-97): pinned by the two-sample closed form, by the second-
moment identity  C = sum_k w_k x_k x_k^T / W - mu mu^T  evaluated in numpy,
and by symmetry and positive semi-definiteness; the per-sky-bin output of a
real estimator run goes through it as well."""
import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth

pytestmark = pytest.mark.gpu


def moments(didj, wgt):
    """x, w, mu, W and the second-moment form of the weighted covariance."""
    on = wgt > 0
    x = np.where(on, didj / np.where(on, wgt, 1.0), 0.0)
    w = np.where(on, wgt, 0.0).sum(axis=1)
    W = w.sum()
    mu = (w[:, None] * x).sum(axis=0) / W
    second = np.einsum("k,ka,kb->ab", w, x, x) / W
    return x, w, mu, W, second - np.outer(mu, mu)


def test_two_samples_closed_form(dev):
    rng = np.random.default_rng(3)
    nb = 23
    x1, x2 = rng.normal(size=nb), rng.normal(size=nb)
    g1, g2 = rng.uniform(1, 5, nb), rng.uniform(1, 5, nb)
    didj = np.stack([x1 * g1, x2 * g2])
    wgt = np.stack([g1, g2])
    cov, mean, W = dev.subsample_covariance(didj, wgt)
    w1, w2 = g1.sum(), g2.sum()
    assert W == pytest.approx(w1 + w2, rel=1e-14)
    # x_k recovered by the division didj / wgt, as on the device
    xs1, xs2 = didj[0] / wgt[0], didj[1] / wgt[1]
    np.testing.assert_allclose(mean, (w1 * xs1 + w2 * xs2) / (w1 + w2), rtol=1e-13, atol=1e-15)
    ref = w1 * w2 / (w1 + w2) ** 2 * np.outer(xs1 - xs2, xs1 - xs2)
    np.testing.assert_allclose(cov, ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("k,nb", [(37, 45), (4, 16), (1, 9), (300, 300), (161, 33)])
def test_against_second_moment_form(dev, k, nb):
    rng = np.random.default_rng(k * 1000 + nb)
    wgt = rng.uniform(0.5, 3.0, (k, nb))
    wgt[rng.random((k, nb)) < 0.2] = 0.0          # empty bins of a sky bin
    didj = rng.normal(size=(k, nb)) * wgt
    print(f"covariance k={k} nbins={nb}", flush=True)
    cov, mean, W = dev.subsample_covariance(didj, wgt)
    x, w, mu, Wr, ref = moments(didj, wgt)
    assert W == pytest.approx(Wr, rel=1e-13)
    np.testing.assert_allclose(mean, mu, rtol=1e-12, atol=1e-14)
    # the second-moment form cancels to about eps * max x^2
    scale = np.abs(ref).max() + 1e-4 * (x ** 2).max()
    assert np.abs(cov - ref).max() <= 1e-11 * scale
    assert np.array_equal(cov, cov.T) or np.abs(cov - cov.T).max() <= 1e-14 * scale
    if k > 1:
        assert np.linalg.eigvalsh((cov + cov.T) / 2).min() >= -1e-10 * scale


def test_no_samples_and_no_weight(dev):
    cov, mean, W = dev.subsample_covariance(np.zeros((0, 7)), np.zeros((0, 7)))
    assert W == 0 and not cov.any() and not mean.any()
    cov, mean, W = dev.subsample_covariance(np.ones((3, 5)), np.zeros((3, 5)))
    assert W == 0 and not cov.any() and not mean.any()


def test_covariance_of_a_per_skybin_run(dev):
    cat = synth.lya_catalog(330, area_deg2=33.0, seed=2001)
    grid = GridSpec.polar()
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4)
    tot, keys, sn, sd, sw = pairs.run_per_skybin(grid, precision=tos_amd.FP64)
    cov, mean, W = dev.subsample_covariance(sd, sw)
    _, _, mu, Wr, ref = moments(sd, sw)
    assert W == pytest.approx(Wr, rel=1e-13)
    scale = np.abs(ref).max()
    assert np.abs(cov - ref).max() <= 1e-11 * scale
    pairs.free()
    dcat.free()
