import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU-only check")


@pytest.fixture(scope="session")
def backend():
    """One HIP backend for the whole GPU session (the library must be built)."""
    from quantumforge_amd.backend import HipBackend

    be = HipBackend(0)
    yield be
    be.synchronize()


def close(a, b, tol):
    a = np.asarray(a)
    b = np.asarray(b)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= tol, f"max abs difference {err:.3e} exceeds {tol:.1e}"
    return err
