"""Seeded synthetic inputs of the benchmark configurations.

* points3d      -- N positions uniform in a cube, delta ~ N(0, 1), w = 1
                   (the 3D brute-force case).
* lya_catalog   -- Lyman-alpha forest sightlines: quasars uniform in area
                   over a sky patch, z_q ~ N(2.4, 0.4) clamped to [2.1, 3.5],
                   pixels on the global grid log10(lambda) = log10(3600) +
                   k * 3e-4 (k >= 1) inside 1040 <= lambda/(1+z_q) <= 1200,
                   delta ~ N(0, 0.2), w = 1/(0.04 + sigma^2) with
                   sigma ~ U(0.05, 0.5); distances from flat LCDM
                   (Omega_m = 0.3) exactly as the h5healxi distance transform
                   does it (h5healxi.cc:196-216), stored as float.

All draws come from numpy's PCG64 with the given seed in a fixed order, so
the same call produces the same arrays on every machine.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

LYA = 1216.0                     # constants.h:10
LOG_LYA = math.log10(LYA)        # constants.h:11
DEG2RAD = math.atan(1.0) * 4 / 180.0


def cosmo_distances(z, omega_lambda=0.7, omega_matter=0.3):
    """Line-of-sight comoving distance and transverse comoving scale (Mpc/h)."""
    z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.float64)
    dc = np.empty_like(z)
    dm = np.empty_like(z)
    _lib.check(_lib.lib().tos_cosmo_distances(float(omega_lambda), float(omega_matter), len(z),
                                              z.ctypes.data, dc.ctypes.data, dm.ctypes.data),
               what="tos_cosmo_distances")
    return dc, dm


@dataclass
class Catalog:
    ra: np.ndarray            # radians, float64
    dec: np.ndarray
    offsets: np.ndarray       # int64, n_los + 1
    value: np.ndarray         # float32 per pixel
    loglam: np.ndarray
    weight: np.ndarray
    distance: np.ndarray
    zq: np.ndarray = None
    scale: float = 0.0        # transverse comoving scale at the smallest redshift
    meta: dict = field(default_factory=dict)

    @property
    def n_los(self) -> int:
        return len(self.ra)

    @property
    def n_pix(self) -> int:
        return int(self.offsets[-1])

    def counts(self) -> np.ndarray:
        return np.diff(self.offsets)

    def subset(self, idx) -> "Catalog":
        idx = np.asarray(idx)
        cnt = self.counts()[idx]
        off = np.zeros(len(idx) + 1, dtype=np.int64)
        np.cumsum(cnt, out=off[1:])
        pix = np.concatenate([np.arange(self.offsets[i], self.offsets[i + 1]) for i in idx]) \
            if len(idx) else np.zeros(0, dtype=np.int64)
        return Catalog(self.ra[idx].copy(), self.dec[idx].copy(), off, self.value[pix].copy(),
                       self.loglam[pix].copy(), self.weight[pix].copy(), self.distance[pix].copy(),
                       None if self.zq is None else self.zq[idx].copy(), self.scale, dict(self.meta))

    def to_c(self):
        """A tos_sightlines struct plus the arrays it points into."""
        keep = [np.ascontiguousarray(self.offsets, dtype=np.int64),
                np.ascontiguousarray(self.ra, dtype=np.float64),
                np.ascontiguousarray(self.dec, dtype=np.float64),
                np.ascontiguousarray(self.value, dtype=np.float32),
                np.ascontiguousarray(self.loglam, dtype=np.float32),
                np.ascontiguousarray(self.weight, dtype=np.float32),
                np.ascontiguousarray(self.distance, dtype=np.float32)]
        s = _lib.Sightlines()
        s.n_los = self.n_los
        s.offsets, s.ra, s.dec, s.value, s.loglam, s.weight, s.distance = [a.ctypes.data for a in keep]
        return s, keep


def patch_half_extent(area_deg2: float) -> tuple:
    """RA extent a and sin(dec) half-range s of a square-ish patch of the given
    area centred on the equator: 2 a s = area and a = 2 asin(s)."""
    area = area_deg2 * DEG2RAD ** 2
    lo, hi = 0.0, 1.0
    for _ in range(200):
        s = 0.5 * (lo + hi)
        if 4.0 * s * math.asin(s) < area:
            lo = s
        else:
            hi = s
    s = 0.5 * (lo + hi)
    a = area / (2.0 * s)
    return a, s


def lya_catalog(n_qso: int, area_deg2: float = 1000.0, seed: int = 2000, *,
                omega_lambda: float = 0.7, omega_matter: float = 0.3,
                lam_min: float = 3600.0, dloglam: float = 3e-4,
                rest_range=(1040.0, 1200.0), z_mean=2.4, z_sigma=0.4, z_clip=(2.1, 3.5),
                delta_sigma=0.2, noise_range=(0.05, 0.5), ra0: float = 0.0) -> Catalog:
    rng = np.random.default_rng(seed)
    a, s = patch_half_extent(area_deg2)
    ra = ra0 + rng.uniform(0.0, a, n_qso)
    dec = np.arcsin(rng.uniform(-s, s, n_qso))
    zq = np.clip(rng.normal(z_mean, z_sigma, n_qso), z_clip[0], z_clip[1])
    # global log-lambda grid, k >= 1 (observed lambda > 3600 A)
    log_min = math.log10(lam_min)
    kmax = int(math.ceil((math.log10(rest_range[1] * (1 + z_clip[1])) - log_min) / dloglam)) + 2
    k = np.arange(1, kmax + 1)
    loglam_k = log_min + k * dloglam
    lam_k = 10.0 ** loglam_k
    lo = rest_range[0] * (1.0 + zq)
    hi = rest_range[1] * (1.0 + zq)
    k_lo = np.searchsorted(lam_k, lo, side="left")
    k_hi = np.searchsorted(lam_k, hi, side="right")
    cnt = np.maximum(k_hi - k_lo, 0)
    offsets = np.zeros(n_qso + 1, dtype=np.int64)
    np.cumsum(cnt, out=offsets[1:])
    npix = int(offsets[-1])
    idx = np.repeat(k_lo - offsets[:-1], cnt) + np.arange(npix)
    loglam = loglam_k[idx].astype(np.float32)
    value = rng.normal(0.0, delta_sigma, npix).astype(np.float32)
    noise = rng.uniform(noise_range[0], noise_range[1], npix)
    weight = (1.0 / (0.04 + noise * noise)).astype(np.float32)
    # distance transform (h5healxi.cc:38-40,200-203): z from the float loglam
    uniq, inv = np.unique(loglam, return_inverse=True)
    zpix = 10.0 ** (uniq.astype(np.float64) - LOG_LYA) - 1.0
    dc, _ = cosmo_distances(zpix, omega_lambda, omega_matter)
    distance = dc.astype(np.float32)[inv]
    cat = Catalog(ra, dec, offsets, value, loglam, weight, distance, zq)
    cat.scale = transverse_scale(cat, omega_lambda, omega_matter)
    cat.meta = dict(n_qso=n_qso, area_deg2=area_deg2, seed=seed)
    return cat


def transverse_scale(cat: Catalog, omega_lambda=0.7, omega_matter=0.3) -> float:
    """h5healxi.cc:194-239: z_min from the smallest FIRST-pixel loglam."""
    first = cat.loglam[cat.offsets[:-1][np.diff(cat.offsets) > 0]]
    zmin = 10.0 ** (float(np.float64(first.min())) - LOG_LYA) - 1.0
    _, dm = cosmo_distances([zmin], omega_lambda, omega_matter)
    return float(dm[0])


def points3d(n: int = 10000, box: float = 100.0, seed: int = 1000):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0.0, box, (3, n))
    d = rng.normal(0.0, 1.0, n)
    w = np.ones(n)
    return xyz[0].copy(), xyz[1].copy(), xyz[2].copy(), d, w


def mock_sizes(n_mocks=128, lo=2000, hi=40000, seed=4000):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, n_mocks)


def job_sizes(n_jobs, lo=1000, hi=100000, seed=5000):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n_jobs)).astype(np.int64)


def plate_tiling(cat: Catalog, spacing_deg: float = 2.5, first_id: int = 3000, seed: int = 7):
    """Plates for PlateBins: centres on a regular (ra, dec) grid of
    `spacing_deg` over the catalogue's patch, ids in shuffled order (so id
    order differs from position order), each sightline on its nearest
    centre.  Returns (plate_of_los, plate_id, plate_ra, plate_dec) as the
    tos_skybins plate arrays expect; plates without sightlines are kept."""
    sp = np.deg2rad(spacing_deg)
    ra0, ra1 = float(cat.ra.min()), float(cat.ra.max())
    de0, de1 = float(cat.dec.min()), float(cat.dec.max())
    gra = np.arange(ra0, ra1 + sp, sp)
    gde = np.arange(de0, de1 + sp, sp)
    cra, cde = [a.ravel() for a in np.meshgrid(gra, gde)]
    ids = first_id + np.random.default_rng(seed).permutation(len(cra)).astype(np.int32)
    # nearest centre by chord distance on the unit sphere
    def unit(ra, dec):
        return np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], axis=-1)
    cu = unit(cra, cde)
    lu = unit(cat.ra, cat.dec)
    near = np.empty(len(cat.ra), dtype=np.int64)
    for s in range(0, len(cat.ra), 4096):
        near[s:s + 4096] = np.argmax(lu[s:s + 4096] @ cu.T, axis=1)
    return ids[near].astype(np.int32), ids, cra, cde
