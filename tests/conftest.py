import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and liblesgo_hip.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def rel_l2(a, b, nx=None):
    """relative L2 error of a against b over the physical columns (x < nx)."""
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = np.asarray(a, dt)
    b = np.asarray(b, dt)
    if nx is not None:
        a = a[..., :nx]
        b = b[..., :nx]
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return num / den if den > 0 else num
