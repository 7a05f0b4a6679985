import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "turbo-octo-spice_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# FP64 sums (both precisions' sum w1 w2 d1 d2, FP64's sum w1 w2) agree with
# the reference's up to summation order: within SUM_RTOL relative.
SUM_RTOL = 1e-10
# A sum of signed terms (sum w1 w2 d1 d2) can cancel to far below the sum of
# its terms' magnitudes, and the rounding error of any summation order is
# relative to those magnitudes, not to the cancelled total.  Such a bin is
# compared relative to CANCEL_FLOOR * sum w1 w2 of the bin when its total is
# smaller than that (with d ~ N(0, 0.2), sum |w1 w2 d1 d2| is about
# 0.025 sum w1 w2, so the floor is 0.4% of the terms' magnitude sum).
CANCEL_FLOOR = 1e-4


def rel_err(a, b):
    """Largest relative difference of a from the reference b over the
    entries where either is nonzero."""
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = (a != 0) | (b != 0)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a - b)[m] / np.maximum(np.abs(b[m]), 1e-300)))


def signed_sum_err(a, b, wgt):
    """rel_err for a per-bin sum of signed terms, with the cancellation
    floor CANCEL_FLOOR * |wgt| (wgt: the reference's sum w1 w2 per bin)."""
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    w = np.abs(np.asarray(wgt, np.float64))
    m = (a != 0) | (b != 0)
    if not m.any():
        return 0.0
    scale = np.maximum(np.maximum(np.abs(b), CANCEL_FLOOR * w), 1e-300)
    return float(np.max(np.abs(a - b)[m] / scale[m]))


def xi_err(a_didj, a_wgt, b_didj, b_wgt):
    """Relative difference of xi = didj / wgt (XiBin::finalize) against the
    reference, with the same cancellation floor (CANCEL_FLOOR in xi)."""
    import numpy as np

    def xi(d, w):
        d = np.asarray(d, np.float64)
        w = np.asarray(w, np.float64)
        m = w > 0
        return np.where(m, d / np.where(m, w, 1.0), 0.0)
    xa, xb = xi(a_didj, a_wgt), xi(b_didj, b_wgt)
    m = (xa != 0) | (xb != 0)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(xa - xb)[m] / np.maximum(np.abs(xb[m]), CANCEL_FLOOR)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")
    config.addinivalue_line("markers", "slow: larger GPU parity cases")


@pytest.fixture(scope="session")
def ora():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def dev():
    import tos_amd
    d = tos_amd.Device(0)
    yield d
    d.close()
