"""The device distance transform (tos_pixel_distances; the per-pixel loop of
h5healxi.cc:196-216) against the host cosmology on the same wavelengths and
against the closed forms of three universes.  The device evaluates
z = 10^(loglam - log10 1216) - 1 and the FP64 integral per pixel; the
synthetic catalogues take z from numpy and the integral from
tos_cosmo_distances (the same arithmetic), so the float distances agree to
within one unit in the last place, nearly always exactly."""
import math

import numpy as np
import pytest

from tos_amd import synth

pytestmark = pytest.mark.gpu

CH = 2997.92458        # c/H0 in Mpc/h


def ulps(a, b):
    return np.abs(np.asarray(a, np.float32).view(np.int32).astype(np.int64)
                  - np.asarray(b, np.float32).view(np.int32).astype(np.int64))


@pytest.mark.parametrize("ol,om", [(0.7, 0.3), (0.73, 0.27), (0.7, 0.0)])
def test_device_transform_equals_host_transform(dev, ol, om):
    cat = synth.lya_catalog(3000, area_deg2=300.0, seed=77, omega_lambda=ol, omega_matter=om)
    print(f"case: {len(cat.loglam)} pixels, Omega_L {ol}, Omega_m {om}", flush=True)
    d = dev.pixel_distances(cat.loglam, ol, om)
    u = ulps(d, cat.distance)
    assert u.max() <= 1
    assert (u == 0).mean() >= 0.999


def test_closed_forms_and_edges(dev):
    loglam = np.linspace(np.log10(3600.0), np.log10(5600.0), 4001).astype(np.float32)
    z = 10.0 ** (loglam.astype(np.float64) - math.log10(1216.0)) - 1.0
    # Einstein-de Sitter: D_C = 2 c/H0 (1 - 1/sqrt(1+z))
    d = dev.pixel_distances(loglam, 0.0, 1.0)
    assert ulps(d, 2 * CH * (1 - 1 / np.sqrt(1 + z))).max() <= 1
    # de Sitter (E = 1): D_C = c/H0 z
    d = dev.pixel_distances(loglam, 1.0, 1e-300)
    assert ulps(d, CH * z).max() <= 1
    # Milne (Omega_k = 1, E = 1 + z): D_C = c/H0 ln(1 + z)
    d = dev.pixel_distances(loglam, 0.0, 1e-300)
    assert ulps(d, CH * np.log1p(z)).max() <= 1
    # z = 0 exactly at log10(1216), and a non-finite wavelength stays NaN
    lz = np.array([math.log10(1216.0), np.nan], np.float32)
    d = dev.pixel_distances(lz, 0.7, 0.3)
    assert abs(float(d[0])) < 1e-3 and math.isnan(float(d[1]))
    assert len(dev.pixel_distances(np.zeros(0, np.float32))) == 0


def test_many_pixels(dev):
    """Two million pixels (several thousand workgroups), against the host
    integral of the distinct wavelengths."""
    rng = np.random.default_rng(5)
    grid = (np.log10(3600.0) + 3e-4 * np.arange(1, 2000)).astype(np.float32)
    loglam = grid[rng.integers(0, len(grid), 2_000_000)]
    d = dev.pixel_distances(loglam, 0.7, 0.3)
    uniq, inv = np.unique(loglam, return_inverse=True)
    dc, _ = synth.cosmo_distances(10.0 ** (uniq.astype(np.float64) - math.log10(1216.0)) - 1.0, 0.7, 0.3)
    u = ulps(d, dc.astype(np.float32)[inv])
    assert u.max() <= 1 and (u == 0).mean() >= 0.999
