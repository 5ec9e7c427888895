"""Test configuration: `gpu` marker, import paths and shared helpers.

CPU tests (-m "not gpu") cover the oracle against the reference's golden
vectors and mathematical pins, the host-side API logic and the C ABI exports.
GPU tests (-m gpu) run the HIP library through the C ABI against the oracle.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hybridsmoothers.jl_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.linalg.norm(b), 1e-300)
    return float(np.linalg.norm(a - b) / den)


@pytest.fixture(scope="session")
def gpu():
    """Initialises device 0 through the C ABI; fails (never skips) on a GPU run."""
    import hybridsmoothers as hs
    hs.init(0)
    return hs
