import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU reference runs")


@pytest.fixture(scope="session")
def gd():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.set_threads(min(8, os.cpu_count() or 1))
    return oracle


def rel_err(a, b):
    """max |a-b| / max |b| (0 if both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.abs(b).max()
    d = np.abs(a - b).max()
    return 0.0 if s == 0 and d == 0 else d / (s if s > 0 else 1.0)


def l2_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())
