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


def assert_at_targets(dt, st, tag=""):
    """every vertex is exactly at the last step's target, except vertices
    whose target another vertex also has (exact duplicates, st["duplicates"]
    of them: one vertex of such a group takes the point); none stalled"""
    pos = dt.download_positions()
    tgt = dt.download_targets()
    off = np.any(pos != tgt, axis=1)
    assert st["stalled"] == 0, "%s: %d vertices stalled" % (tag, st["stalled"])
    if off.any():
        # each off-target vertex shares its target with another vertex
        keys = tgt.view(np.uint64).ravel()
        _, inv, cnt = np.unique(keys, return_inverse=True, return_counts=True)
        assert np.all(cnt[inv[off]] >= 2), "%s: off-target vertices %s" % (tag, np.flatnonzero(off & (cnt[inv] < 2))[:8])
    assert off.sum() == st["duplicates"], "%s: %d off target, %d duplicates" % (tag, off.sum(), st["duplicates"])
    return pos
