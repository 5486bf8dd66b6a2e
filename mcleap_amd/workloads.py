"""Synthetic workloads of the benchmark configurations.

  c1  box,   N = 10,000,    Gaussian sigma = 1e-3, clamped, FRNN r = 0.02
  c2  torus, N = 2^20,      shear (1e-3 sin 2 pi y, 1e-3 cos 2 pi x) + sigma 2e-4,
                            wrapped, FRNN r = 2/sqrt(N) on a periodic grid
  c3  box,   N = 2^22,      64 Gaussian blobs (sigma 0.02), Brownian sigma = 0.5/sqrt(N),
                            FRNN r = 3/sqrt(N)
  c4  box,   1000 x 1000 lattice k/1000 + U(+-1e-7) in FP32, sigma = 1e-7
  c5  box,   4096 instances, n_i = floor(exp U[ln 256, ln 65536]), sigma = 5e-4,
             FRNN r_i = 2/sqrt(n_i)
Initial points come from numpy's PCG64 with the config's seed; duplicate
points are dropped (the triangulation rejects them).  Per-step motion is
generated on the GPU by a counter-based Philox4x32-10 stream keyed by the
seed (mcl_step_params_t).
"""
import numpy as np

from . import _lib

CONFIGS = {
    "c1": dict(n=10_000, seed=1, domain="box", motion=_lib.MOTION_GAUSS, sigma=1e-3, amp=0.0, steps=50,
               radius=0.02),
    "c2": dict(n=1 << 20, seed=2, domain="torus", motion=_lib.MOTION_SHEAR, sigma=2e-4, amp=1e-3, steps=200,
               radius=2.0 / 1024.0),
    "c3": dict(n=1 << 22, seed=3, domain="box", motion=_lib.MOTION_GAUSS, sigma=0.5 / 2048.0, amp=0.0, steps=100,
               radius=3.0 / 2048.0),
    "c4": dict(n=1_000_000, seed=4, domain="box", motion=_lib.MOTION_GAUSS, sigma=1e-7, amp=0.0, steps=50,
               radius=None),
}


def unique_points(xy):
    _, idx = np.unique(xy, axis=0, return_index=True)
    return xy[np.sort(idx)]


def initial_points(name, n=None):
    cfg = CONFIGS[name]
    n = n or cfg["n"]
    rng = np.random.default_rng(cfg["seed"])
    if name in ("c1", "c2"):
        xy = rng.random((n, 2), dtype=np.float64).astype(np.float32)
        xy[xy >= 1.0] = 0.0
    elif name == "c3":
        centres = 0.1 + 0.8 * rng.random((64, 2))
        which = rng.integers(0, 64, n)
        xy = centres[which] + 0.02 * rng.standard_normal((n, 2))
        xy = np.clip(xy, 0.0, 1.0 - 2.0 ** -24).astype(np.float32)
    elif name == "c4":
        side = int(round(np.sqrt(n)))
        g = np.stack(np.meshgrid(np.arange(side), np.arange(side), indexing="ij"), -1).reshape(-1, 2) / float(side)
        xy = (g + rng.uniform(-1e-7, 1e-7, g.shape)).astype(np.float32)
        xy = np.clip(xy, 0.0, 1.0 - 2.0 ** -24).astype(np.float32)
    else:
        raise KeyError(name)
    return unique_points(xy)


def grid_shape(radius, domain):
    """cells of side >= radius covering [0, 1)"""
    nc = max(3, int(np.floor(1.0 / radius)))
    return nc, 1.0 / nc
