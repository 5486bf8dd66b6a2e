"""Autotuner: measures, on the GPU it runs on, the timestep of each benchmark
workload under candidate launch and round-loop parameters and writes the
fastest set per workload class to a table that libmcleap_hip.so loads when a
triangulation or grid is created (tuning.json next to the library).

    python -m mcleap_amd.autotune [--workloads c2,c1,c4,c5] [--out PATH]

The reference fixes these by hand (blocksize = 128 in
Device_Triangulation.h:151, blocksize = 64 in gridCount2d.h:35, cells of the
query radius).  Parameters swept, one at a time from the defaults
(coordinate descent; a candidate is kept when it is faster by more than 2%
than the current choice measured right before and right after it):
  grid_block             threads per block of the grid sort and FRNN kernels
  frnn_cells_per_radius  FRNN cells per query radius (1: 3 x 3 cells; 2: 5 x 5 half cells)
  graph                  hipGraph replay of the star step's head and the tail
  star_head_passes       star passes before the first host read
  tail_threshold         Lawson queue length finished by one workgroup
  revert_tail, reloc_tail  the same for take-backs and relocations
  big_grid_min           big-exact Lawson queue length run on 256 workgroups
Each measurement is the mean time of whole timesteps (kinetic step, one-ring
CSR, grid sort and FRNN query; about 0.3 s of them) after `warmup` more, on
the same triangulation, so the motion carries on between measurements.  At
the end the tuned set and the defaults are measured alternately twice; the
defaults are written unless the tuned set is faster by more than 1%.
"""
import argparse
import json
import os
import sys
import time

os.environ.setdefault("MCL_TUNING", "off")  # measure from the library's defaults

import numpy as np  # noqa: E402

from . import _lib  # noqa: E402
from .workloads import CONFIGS, ensemble_points, grid_shape, initial_points  # noqa: E402

CANDIDATES = {
    "grid_block": [256, 128, 64],
    "frnn_cells_per_radius": [1, 2],
    "graph": [0, 1],
    "star_head_passes": [4, 3, 5, 2, 6],
    "tail_threshold": [8192, 2048, 4096, 16384, 32768],
    "revert_tail": [1024, 256, 4096],
    "reloc_tail": [1024, 256, 4096],
    "big_grid_min": [2048, 512, 8192],
}
STAR_PARAMS = ["graph", "star_head_passes"]
FLIP_PARAMS = ["tail_threshold", "revert_tail", "reloc_tail", "big_grid_min"]


class Bench:
    """one workload's triangulation, grid and buffers, stepped on demand"""

    def __init__(self, name, ninst=None):
        import torch

        import mcleap_amd as M
        self.M, self.torch = M, torch
        self.cfg = dict(CONFIGS[name])
        self.name = name
        self.batch = name == "c5"
        periodic = self.cfg["domain"] == "torus"
        self.domain = "torus" if periodic else ("ensemble" if self.batch else "box")
        if self.batch:
            pts = ensemble_points(ninst)
            self.dt = M.DeviceTriangulation.from_ensemble(pts)
            H = self.dt.h_triangulation
            rad_i = np.array([2.0 / np.sqrt(max(int(c), 1)) for c in H.counts], np.float32)
            self.bgrid = M.GridCount2dBatch(H.vstart, np.minimum(rad_i, np.float32(1.0 / 3.0)))
            self.n = int(sum(len(p) for p in pts))
        else:
            xy = initial_points(name)
            self.dt = M.DeviceTriangulation.from_points(xy, periodic=periodic)
            self.bgrid = None
            self.n = len(xy)
        nv = self.dt.m_pos_size
        self.nv = nv
        self.ring = (torch.zeros(nv + 1, dtype=torch.int32, device="cuda"),
                     torch.zeros(self.dt.m_he_size, dtype=torch.int32, device="cuda"))
        self.f_off = torch.zeros((self.bgrid.numP if self.batch else nv) + 1, dtype=torch.int32, device="cuda")
        self.f_nbr = None
        self.grid = None
        self.has_grid = self.batch or bool(self.cfg["radius"])
        self.params = {}
        self.k = 0
        if self.has_grid:
            self.set_grid(256, 1)

    def set_grid(self, block, cells):
        """a grid with cells of radius / cells (the reference's: cells of the radius)"""
        M, torch = self.M, self.torch
        if self.batch:
            self.bgrid.build(self.dt.pos_tensor())
            tot = self.bgrid.frnn_csr(self.f_off)
        else:
            r = self.cfg["radius"]
            nc, cs = grid_shape(r / cells, self.cfg["domain"])
            flags = M.GRID_MORTON | (M.GRID_PERIODIC if self.cfg["domain"] == "torus" else 0)
            self.grid = M.GridCount2d(self.nv, (0.0, 0.0), (cs, cs), (nc, nc), flags)
            self.grid.set_block(block)
            self.grid.build(self.dt.pos_tensor())
            tot = self.grid.frnn_csr(self.grid.sorted_ptr(), r, self.f_off)
        self.f_nbr = torch.zeros(int(tot * 1.25) + 1024, dtype=torch.int32, device="cuda")

    def set(self, key, value):
        if key in ("grid_block", "frnn_cells_per_radius"):
            self.params[key] = value
            if not self.batch and self.has_grid:
                self.set_grid(self.params.get("grid_block", 256), self.params.get("frnn_cells_per_radius", 1))
            return
        self.params[key] = value
        self.dt.set_option(_lib.TRI_OPTIONS[key], int(value))

    def step(self):
        c = self.cfg
        if self.batch:
            self.dt.step(c["motion"], c["sigma"], c["amp"], seed=c["seed"], step=self.k)
            self.dt.one_ring_csr_async(*self.ring)
            self.bgrid.build(self.dt.pos_tensor())
            self.bgrid.frnn_csr_async(self.f_off, self.f_nbr)
        else:
            self.dt.timestep(c["motion"], c["sigma"], c["amp"], seed=c["seed"], step=self.k, ring=self.ring,
                             grid=self.grid, rad=c["radius"] or 0.0,
                             frnn=(self.f_off, self.f_nbr) if self.grid is not None else None)
        self.k += 1

    def measure(self, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / steps


def tune(name, steps, warmup, ninst=None, budget_ms=300.0, log=print):
    b = Bench(name, ninst)
    star = b.domain == "torus"
    # each workload is tuned on the path the step takes for it (the torus by
    # star recomputation, the box workloads by flips), with the path choice's
    # probing of the other path switched off while measuring
    b.dt.set_option(_lib.TRI_OPTIONS["star_enable"], 1 if star else 0)
    for _ in range(5):
        b.step()
    params = ["grid_block", "frnn_cells_per_radius"] if (b.has_grid and not b.batch) else []
    params += STAR_PARAMS if star else FLIP_PARAMS
    best = {p: CANDIDATES[p][0] for p in params}
    for p, v in best.items():
        b.set(p, v)
    default_ms = b.measure(steps, warmup)
    # enough steps per measurement for about budget_ms of work
    steps = int(min(200, max(steps, budget_ms / max(default_ms, 1e-3))))
    default_ms = b.measure(steps, warmup)
    log("[tune] %s: default %.3f ms/step" % (name, default_ms))
    # the motion changes the mesh as the measurements go on, so each
    # candidate is compared with the current choice measured right before and
    # right after it
    for p in params:
        for v in CANDIDATES[p][1:]:
            ref0 = b.measure(steps, warmup)
            b.set(p, v)
            ms = b.measure(steps, warmup)
            b.set(p, best[p])
            ref1 = b.measure(steps, warmup)
            ref = 0.5 * (ref0 + ref1)
            keep = ms < 0.98 * ref
            log("[tune] %s: %s = %s -> %.3f ms/step against %.3f%s" % (name, p, v, ms, ref, " (kept)" if keep else ""))
            if keep:
                best[p] = v
                b.set(p, v)
    # the tuned set against the defaults, alternating (A B A B), so that a set
    # chosen on noise is not written
    default = {p: CANDIDATES[p][0] for p in params}
    t_tuned, t_default = [], []
    for _ in range(2):
        for p in params:
            b.set(p, default[p])
        t_default.append(b.measure(steps, warmup))
        for p in params:
            b.set(p, best[p])
        t_tuned.append(b.measure(steps, warmup))
    tuned_ms, default_again = float(np.mean(t_tuned)), float(np.mean(t_default))
    kept = tuned_ms < 0.99 * default_again
    if not kept:
        best = default
    b.dt.set_option(_lib.TRI_OPTIONS["star_enable"], 2)
    log("[tune] %s: tuned %.3f ms/step against the defaults %.3f: %s (%d steps per measurement)"
        % (name, tuned_ms, default_again, "tuned set written" if kept else "defaults written", steps))
    entry = {"domain": b.domain, "log2_nv": int(round(np.log2(b.nv))), "workload": name, "vertices": b.nv}
    entry.update({p: int(v) for p, v in best.items()})
    entry.update({"ms_per_step_default": round(default_again, 4), "ms_per_step_tuned": round(tuned_ms, 4),
                  "tuned_set_kept": bool(kept), "steps": steps})
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="c2,c1,c4,c5")
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ninst", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json"))
    a = ap.parse_args()
    import torch
    entries = []
    for w in a.workloads.split(","):
        entries.append(tune(w.strip(), a.steps, a.warmup, a.ninst or None, log=lambda s: print(s, flush=True)))
        torch.cuda.empty_cache()
    out = {"generator": "mcleap_amd/autotune.py", "device": torch.cuda.get_device_name(0),
           "library": _lib.lib().mcl_version().decode(), "steps_per_measurement": a.steps,
           "entries": entries}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
