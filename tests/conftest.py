import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: larger CPU oracle runs")


@pytest.fixture(scope="session")
def oracle():
    """The CPU parity oracle (restatement of the reference's cpu:: path)."""
    path = os.path.join(ROOT, "oracle", "_build")
    if path not in sys.path:
        sys.path.insert(0, path)
    import gprat_oracle
    return gprat_oracle


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(GOLDEN, "data_1024")


def within_rel(a, b, tol):
    """Catch2 WithinRel: |a-b| <= tol * max(|a|, |b|) element-wise."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) <= tol * np.maximum(np.abs(a), np.abs(b))


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    den = np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b) / den))


def norm_rel(a, b):
    """||a - b||_inf / ||b||_inf."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
