"""Full-survey size (160,000 quasars over 10,000 deg^2, ~2.3e7 pixels, the
precision-test workload): properties that do not need the CPU oracle.

  * FP32C bins every pixel pair exactly as FP64: counts equal bin by bin,
    sums within 1e-10 (both add the reference's FP32 terms in FP64).
  * The counts add up to the pixel pairs the pass reports as used.
  * One histogram per sky bin sums to the plain pass.
  * The GPU candidate search gives the host search's pair list.
"""
import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, synth
from conftest import SUM_RTOL, rel_err, signed_sum_err, xi_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def survey():
    return synth.lya_catalog(160000, area_deg2=10000.0, seed=3000)


@pytest.mark.parametrize("gname", ["cartesian", "polar"])
def test_fp32c_bins_equal_fp64_at_full_size(dev, survey, gname):
    grid = getattr(GridSpec, gname)()
    print(f"full size {gname}", flush=True)
    dcat = dev.upload(survey)
    pairs = dcat.pairs(grid.max_angular_scale(survey.scale), order=4, nthreads=16, on_device=True)
    h64 = pairs.run(grid, precision=tos_amd.FP64)
    h32 = pairs.run(grid, precision=tos_amd.FP32C)
    assert np.array_equal(h32.num_pairs, h64.num_pairs)
    assert signed_sum_err(h32.didj, h64.didj, h64.wgt) <= SUM_RTOL
    assert rel_err(h32.wgt, h64.wgt) <= 1e-10
    assert int(h64.num_pairs.sum()) == h64.num_pixel_pairs_used
    assert h64.num_pairs_evaluated <= pairs.stats()["num_pixel_pairs"]
    # per-sky-bin histograms add up to the plain pass
    tot, keys, sn, sd, sw = pairs.run_per_skybin(grid, precision=tos_amd.FP32C)
    assert np.array_equal(tot.num_pairs, h32.num_pairs)
    assert np.array_equal(sn.sum(axis=0), h32.num_pairs)
    assert rel_err(sw.sum(axis=0), h32.wgt) <= 1e-10
    pairs.free()
    dcat.free()


def test_device_search_equals_host_at_full_size(dev, survey):
    grid = GridSpec.cartesian()
    dcat = dev.upload(survey)
    ma = grid.max_angular_scale(survey.scale)
    a = dcat.pairs(ma, order=4, nthreads=16, on_device=True)
    b = dcat.pairs(ma, order=4, nthreads=16, on_device=False)
    assert a.stats() == b.stats()
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    a.free()
    b.free()
    dcat.free()
