import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libmcleap_hip.so")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


def uniform_points(n, seed, lo=0.0, hi=1.0):
    """seeded uniform FP32 points, duplicates removed (keeps first)"""
    r = np.random.default_rng(seed)
    xy = (lo + (hi - lo) * r.random((n, 2))).astype(np.float32)
    _, idx = np.unique(xy, axis=0, return_index=True)
    return xy[np.sort(idx)]
