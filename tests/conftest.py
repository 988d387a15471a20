import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available():
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def rel_err(a, b, floor=0.0):
    """max |a - b| / max(max |b|, floor) over the whole array (norm-wise
    relative error; `floor` keeps fields that are pure round-off noise, such
    as the plasma response ahead of the beam, from dividing noise by noise)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)) if b.size else 0.0, floor)
    diff = np.max(np.abs(a - b)) if b.size else 0.0
    if scale == 0.0:
        return float(diff)
    return float(diff / scale)
