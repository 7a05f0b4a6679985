"""HDF5 delta files (HDF5Delta.cc:40-93): the writer/loader round trip and
the loader's filtering, on the CPU.  The C++ tools link the same
libturbooctospice.so the h5healxi driver uses."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def cat():
    from tos_amd import synth
    return synth.lya_catalog(40, area_deg2=4.0, seed=31)


def test_round_trip(tmp_path, cat):
    from tos_amd import h5io
    path = str(tmp_path / "delta.h5")
    h5io.write_h5(cat, path, plates=np.arange(cat.n_los) % 7 + 100)
    back, plate = h5io.read_h5(path)
    assert back.n_los == cat.n_los
    assert np.array_equal(back.offsets, cat.offsets)
    for f in ("ra", "dec"):
        assert np.array_equal(getattr(back, f), getattr(cat, f))
    for f in ("value", "loglam", "weight", "distance"):
        assert np.array_equal(getattr(back, f), getattr(cat, f))
    assert np.array_equal(plate, np.arange(cat.n_los) % 7 + 100)


def test_loader_filters_pixels_and_empty_sightlines(tmp_path, cat):
    """Pixels with non-finite delta or weight, or weight <= 0, are dropped;
    a sightline left empty is skipped (HDF5Delta.cc:74,80-83)."""
    from tos_amd import h5io
    v = cat.value.copy()
    w = cat.weight.copy()
    o = cat.offsets
    v[o[0]] = np.nan                 # LOS 0: one NaN delta
    w[o[1] + 1] = 0.0                # LOS 1: one zero weight
    w[o[2] + 2] = -1.0               # LOS 2: one negative weight
    w[o[3] + 3] = np.inf             # LOS 3: one infinite weight
    w[o[4]:o[5]] = 0.0               # LOS 4: every weight zero -> skipped
    path = str(tmp_path / "filtered.h5")
    h5io.write_h5(cat, path, value=v, weight=w)
    back, _ = h5io.read_h5(path)
    assert back.n_los == cat.n_los - 1
    counts = np.diff(back.offsets)
    expect = np.delete(cat.counts(), 4)
    expect[:4] -= 1
    assert np.array_equal(counts, expect)
    # the sixth stored sightline became the fifth loaded one
    assert back.ra[4] == cat.ra[5]
    keep = np.isfinite(v) & np.isfinite(w) & (w > 0)
    assert np.array_equal(back.value, cat.value[keep])
    assert np.array_equal(back.weight, w[keep])
