"""CPU checks of the C-ABI library: it loads without a GPU, exports every
entry point include/tos_amd.h declares, and its host-only functions (grid
parsing, bin centres, cosmology distances, HEALPix) are right.  No compute
entry point is called here."""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tos_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tos_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("tos_ctx_create", "tos_catalog_upload", "tos_pairs_build", "tos_xi_run",
                 "tos_batch_run", "tos_server_start", "tos_job_submit", "tos_brute3d"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import ctypes
    from tos_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the whole surface too
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_grid_parse_matches_python():
    import ctypes
    from tos_amd import _lib, GridSpec
    L = _lib.lib()
    for spec in ("[0:200]*50", "[0:1]*20", "[3.556:3.735]*1", "[-5:5]*7", "[0:0.06]*6"):
        g = _lib.Grid()
        assert L.tos_grid_parse_axis(spec.encode(), ctypes.byref(g), 1) == 0
        n, lo, w, hi = __import__("tos_amd").parse_binning(spec)
        assert (g.nbins[1], g.min[1], g.width[1], g.max[1]) == (n, lo, w, hi)
    g = _lib.Grid()
    for bad in ("[1:0]*5", "[0:1]*0", "0:1*5", "[0:1]"):
        assert L.tos_grid_parse_axis(bad.encode(), ctypes.byref(g), 0) != 0


def test_max_angular_scale_formulas():
    from tos_amd import GridSpec
    scale = 3587.75
    assert GridSpec.cartesian().max_angular_scale(scale) == 200.0 / scale
    assert GridSpec.polar(mu="[0.5:1]*10").max_angular_scale(scale) == 200.0 * math.sqrt(1 - 0.25) / scale
    assert GridSpec.observing(theta="[0:0.06]*6").max_angular_scale(scale) == 0.06


def test_bin_centers_row_major():
    import ctypes
    from tos_amd import _lib, GridSpec
    g = GridSpec.cartesian(z="[3.556:3.735]*2").to_c()
    c = (ctypes.c_double * 3)()
    # index = (i0 * n1 + i1) * n2 + i2
    assert _lib.lib().tos_grid_bin_centers(ctypes.byref(g), (3 * 50 + 7) * 2 + 1, c) == 0
    assert (c[0], c[1]) == (14.0, 30.0)
    assert c[2] == pytest.approx(3.556 + 1.5 * 0.0895)


def test_cosmology_closed_forms():
    from tos_amd.synth import cosmo_distances
    z = np.array([0.1, 0.5, 1.0, 2.4, 3.5])
    ch = 2997.92458
    # Einstein-de Sitter: D_C = 2 c/H0 (1 - 1/sqrt(1+z))
    dc, dm = cosmo_distances(z, omega_lambda=0.0, omega_matter=1.0)
    np.testing.assert_allclose(dc, 2 * ch * (1 - 1 / np.sqrt(1 + z)), rtol=1e-12)
    np.testing.assert_allclose(dm, dc, rtol=0)
    # de Sitter (Omega_L = 1): D_C = c/H0 z
    dc, _ = cosmo_distances(z, omega_lambda=1.0, omega_matter=1e-300)
    np.testing.assert_allclose(dc, ch * z, rtol=1e-12)
    # flat LCDM against composite Simpson with 2e5 panels
    dc, _ = cosmo_distances(z, omega_lambda=0.7, omega_matter=0.3)
    for zi, di in zip(z, dc):
        x = np.linspace(0, zi, 200001)
        f = 1 / np.sqrt(0.3 * (1 + x) ** 3 + 0.7)
        h = x[1] - x[0]
        simpson = h / 3 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
        assert di == pytest.approx(ch * simpson, rel=1e-12)
    # open universe: D_M = c/H0 sinh(sqrt(Ok) D_C H0/c) / sqrt(Ok)
    dc, dm = cosmo_distances(z, omega_lambda=0.0, omega_matter=0.3)
    ok = 0.7
    np.testing.assert_allclose(dm, ch * np.sinh(np.sqrt(ok) * dc / ch) / np.sqrt(ok), rtol=1e-13)


def test_omega_matter_zero_means_flat():
    from tos_amd.synth import cosmo_distances
    a, _ = cosmo_distances([2.0], omega_lambda=0.7, omega_matter=0.3)
    from tos_amd import _lib
    import ctypes
    out = np.zeros(1)
    zz = np.array([2.0])
    assert _lib.lib().tos_cosmo_distances(0.7, 0.0, 1, zz.ctypes.data, out.ctypes.data, None) == 0
    assert out[0] == a[0]


def test_synthetic_catalog_shape():
    from tos_amd import synth
    cat = synth.lya_catalog(2000, area_deg2=200.0, seed=2000)
    n = cat.counts()
    assert n.min() >= 47 and n.max() <= 208
    assert 140 < n.mean() < 152
    assert cat.zq.min() >= 2.1 and cat.zq.max() <= 3.5
    lam = 10 ** cat.loglam.astype(np.float64)
    assert lam.min() > 3600.0
    w = cat.weight
    assert w.min() >= 1 / (0.04 + 0.25) - 1e-3 and w.max() <= 1 / (0.04 + 0.0025) + 1e-3
    # distance grows with wavelength inside every sightline
    for k in range(20):
        d = cat.distance[cat.offsets[k]:cat.offsets[k + 1]]
        assert np.all(np.diff(d) > 0)
    # same seed, same arrays
    again = synth.lya_catalog(2000, area_deg2=200.0, seed=2000)
    assert np.array_equal(again.distance, cat.distance) and np.array_equal(again.ra, cat.ra)


# ---------------------------------------------------------------------------
# Covariance diagnostics (XiEstimator.cc:82-87): positive definiteness and
# log|C| from the host Cholesky, against numpy's LU-based slogdet and the
# closed forms of diagonal and block matrices.
# ---------------------------------------------------------------------------
def test_covariance_check_spd_matches_slogdet():
    import tos_amd
    rng = np.random.default_rng(77)
    for n in (1, 3, 64, 65, 200):
        x = rng.normal(size=(n + 40, n))
        c = x.T @ x / (n + 40) + 0.05 * np.eye(n)
        pd, ld = tos_amd.covariance_check(c)
        sign, ref = np.linalg.slogdet(c)
        assert pd and sign > 0
        assert abs(ld - ref) <= 1e-9 * max(1.0, abs(ref)), (n, ld, ref)


def test_covariance_check_closed_forms():
    import tos_amd
    d = np.linspace(0.5, 3.0, 150)
    pd, ld = tos_amd.covariance_check(np.diag(d))
    assert pd and abs(ld - np.log(d).sum()) <= 1e-12 * abs(np.log(d).sum())
    # [[a, b], [b, a]] blocks: det = a^2 - b^2 each
    a, b = 2.0, 1.5
    c = np.kron(np.eye(40), np.array([[a, b], [b, a]]))
    pd, ld = tos_amd.covariance_check(c)
    assert pd and abs(ld - 40 * math.log(a * a - b * b)) <= 1e-11
    # only the lower triangle is read
    u = np.triu(np.full((40, 40), 1e30), 1)
    pd2, ld2 = tos_amd.covariance_check(np.diag(d[:40]) + u)
    assert pd2 and abs(ld2 - np.log(d[:40]).sum()) <= 1e-12 * abs(np.log(d[:40]).sum())


def test_covariance_check_not_positive_definite():
    import tos_amd
    rng = np.random.default_rng(78)
    q, _ = np.linalg.qr(rng.normal(size=(90, 90)))
    ev = np.linspace(0.1, 2.0, 90)
    ev[70] = -0.3                                  # one negative eigenvalue
    pd, ld = tos_amd.covariance_check(q @ np.diag(ev) @ q.T)
    assert not pd and math.isnan(ld)
    # an empty bin: a zero row and column (pivot exactly 0)
    c = np.diag(np.linspace(1.0, 2.0, 100))
    c[57, 57] = 0.0
    assert not tos_amd.covariance_check(c)[0]
    # a non-finite entry
    c = np.eye(10)
    c[4, 2] = np.nan
    assert not tos_amd.covariance_check(c)[0]
    with pytest.raises(ValueError):
        tos_amd.covariance_check(np.zeros((3, 4)))


def test_cosmology_against_adaptive_quadrature():
    """The FP64 line-of-sight distance (own exp / log1p, shared with the
    device transform) against scipy's adaptive quadrature of 1/E(z) at
    1e-13, flat, open and closed universes, z from 1e-6 to 6."""
    from scipy import integrate
    from tos_amd.synth import cosmo_distances
    z = np.array([1e-6, 1e-3, 0.1, 1.0, 2.0, 3.5, 6.0])
    for ol, om in ((0.7, 0.3), (0.73, 0.27), (0.0, 1.0), (0.3, 0.3), (0.9, 0.3)):
        dc, _ = cosmo_distances(z, ol, om)
        ok = 1.0 - om - ol
        for zi, di in zip(z, dc):
            ref, _ = integrate.quad(lambda x: 1.0 / math.sqrt(om * (1 + x) ** 3 + ok * (1 + x) ** 2 + ol), 0.0, zi,
                                    epsabs=0.0, epsrel=1e-13, limit=200)
            assert abs(di - 2997.92458 * ref) <= 1e-13 * 2997.92458 * ref, (ol, om, zi)


def test_new_entry_points_refuse_invalid_arguments():
    """Argument checks of the C-ABI that run before any device work."""
    import ctypes
    from tos_amd import _lib
    L = _lib.lib()
    ERR_INVALID = 1     # TOS_ERR_INVALID
    x = np.zeros(4, np.float32)
    assert L.tos_pixel_distances(None, 0.7, 0.3, 4, x.ctypes.data, x.ctypes.data) == ERR_INVALID
    pd, ld = ctypes.c_int32(0), ctypes.c_double(0)
    c = np.eye(3)
    assert L.tos_covariance_check(0, c.ctypes.data, ctypes.byref(pd), ctypes.byref(ld)) == ERR_INVALID
    assert L.tos_covariance_check(3, None, ctypes.byref(pd), ctypes.byref(ld)) == ERR_INVALID
    shape = (ctypes.c_int32 * 4)()
    assert L.tos_xi_tuned_shape(None, None, 1, ctypes.byref(shape)) == ERR_INVALID
    assert L.tos_xi_run_noalloc(None, None, None, None, 1, None, None, None, None) == ERR_INVALID
