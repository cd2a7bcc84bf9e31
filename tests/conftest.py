import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


_HIER = {}


def sa_hierarchy(grid):
    """Smoothed-aggregation hierarchy of the Poisson operator on `grid` (cached)."""
    key = ("sa", tuple(grid))
    if key not in _HIER:
        from amjax_amd import gallery, hierarchy
        _HIER[key] = hierarchy.smoothed_aggregation_solver(gallery.poisson(grid))
    return _HIER[key]


def rs_hierarchy(grid):
    key = ("rs", tuple(grid))
    if key not in _HIER:
        from amjax_amd import gallery, hierarchy
        _HIER[key] = hierarchy.ruge_stuben_solver(gallery.poisson(grid))
    return _HIER[key]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a ROCm GPU")
    return torch.device("cuda", 0)
