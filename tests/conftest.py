import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REFERENCE_BIN = "/root/reference/input_files/icos2562_tc5_input.bin"
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def pkg():
    import swe_pkg
    return swe_pkg.load()


@pytest.fixture(scope="session")
def pyoracle():
    import pyoracle as po
    return po


def have_reference() -> bool:
    return os.path.exists(REFERENCE_BIN)


def rel_by_var(got, want, prob=None):
    """Per-variable relative max-norm error, (u, v, w, h)."""
    g = np.asarray(got)
    w = np.asarray(want)
    if prob is not None:
        g = prob.to_node_major(g)
        w = prob.to_node_major(w)
    else:
        g = g.reshape(-1, 4)
        w = w.reshape(-1, 4)
    num = np.abs(g - w).max(axis=0)
    den = np.abs(w).max(axis=0)
    return num / np.where(den > 0, den, 1.0)
