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
    "c5": dict(ninst=4096, seed=5, domain="box", motion=_lib.MOTION_GAUSS, sigma=5e-4, amp=0.0, steps=100,
               radius=None, nmin=256, nmax=65536),
}


def unique_points(xy):
    """xy without repeated points (first occurrence kept, order preserved)"""
    b = np.ascontiguousarray(xy + np.float32(0.0), np.float32).view(np.uint32).astype(np.uint64)  # -0 -> +0
    key = (b[:, 0] << np.uint64(32)) | b[:, 1]
    _, idx = np.unique(key, return_index=True)
    if len(idx) == len(xy):
        return xy
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


def ensemble_sizes(ninst=None, seed=None, nmin=None, nmax=None):
    """instance sizes of the batched ensemble: floor(exp U[ln nmin, ln nmax]),
    by default nmin = 256, nmax = 65536"""
    cfg = CONFIGS["c5"]
    rng = np.random.default_rng(cfg["seed"] if seed is None else seed)
    u = rng.uniform(np.log(nmin or cfg["nmin"]), np.log(nmax or cfg["nmax"]), ninst or cfg["ninst"])
    return np.floor(np.exp(u)).astype(np.int64)


def ensemble_points(ninst=None, seed=None, nmin=None, nmax=None, sizes=None):
    """one (n_i, 2) float32 array per instance, uniform in [0, 1)^2, from
    per-instance PCG64 streams (seed, i), duplicates dropped"""
    cfg = CONFIGS["c5"]
    seed = cfg["seed"] if seed is None else seed
    if sizes is None:
        sizes = ensemble_sizes(ninst, seed, nmin, nmax)
    out = []
    for i, n in enumerate(sizes):
        rng = np.random.default_rng([seed, i])
        xy = rng.random((int(n), 2), dtype=np.float64).astype(np.float32)
        xy[xy >= 1.0] = 0.0
        out.append(unique_points(xy))
    return out
