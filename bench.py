"""Kinetic Delaunay benchmark on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload c2] [--gpus 1]

A step is one pass of the hot path over the whole point set: move every
point, restore the exact Delaunay triangulation (on the torus: per-vertex
star recomputation with consistency checks; in the box: repair flips /
take-backs / relocation / Lawson flips), one-ring CSR, and the uniform-grid
FRNN rebuild + query (CSR).  Inputs live in HBM before the timed region; the
K timed steps are bracketed by torch.cuda.synchronize().  After the timed
steps the final mesh is read through the half-edge arrays and checked (every
edge exactly locally Delaunay, no inverted triangle).  Prints one JSON line.

roofline: the dominant kernel class (largest GPU time) among the library's
per-kernel HIP event timers, credited with algorithmic bytes per launch from
the per-unit model (per point: move 16 B, inversion check 32 B, Delaunay
check 56 B, star 390 B, grid build 43 B, FRNN count 13 B, FRNN fill 4*k B;
per relocated vertex: 648 B).  traffic: the HBM bytes per launch of that
kernel measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a copy
kernel) for this workload, read from profiles/pmc_traffic.json when it holds
them.  valu_roofline: the same launch's VALU instruction count (SQ_INSTS_VALU,
same file) over the measured launch time against the issue peak (256 CUs x 4
SIMDs x one wave64 instruction per 4 cycles at 2.4 GHz).  step_roofline: the
whole step credited with (212 + 4k) B per point.
cpu_baseline: the oracle's exact host triangulation rebuilt after each move,
one thread.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
PEAK_VALU_WAVE_INST = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second
BYTES = {"move": 16, "check": 32, "lawson": 56, "grid_build": 43, "frnn_count": 13}
RELOC_BYTES = 648  # per relocated vertex: star read 3*6*16 + 7 positions * 8 + rewrite 3*6*16
STAR_BYTES = 390   # per vertex: own star 28 + 6 neighbour stars 6*28 + ~19 candidate positions 152 + new star 28 + pos 8


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--n", type=int, default=0, help="override point count (testing only)")
    ap.add_argument("--ninst", type=int, default=0, help="override the ensemble's instance count (testing only)")
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--build", choices=("device", "host"), default="device",
                    help="initial triangulation on the GPU (mcl_tri_build, mcl_batch_build) or on the host")
    ap.add_argument("--plain-steps", type=int, default=10, help="steps of the plain form timed for comparison")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (reference for HBM byte counters)")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy of known size: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    its counters give the factor between counted and actual bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}))


def cpu_baseline(pos0, cfg, periodic, seconds):
    """The oracle (restated HostTriangulation: walk, f1to3, delonizeVertex,
    exact predicates) rebuilding the exact Delaunay triangulation of the moved
    points each step, one thread.  The reference's own kinetic host path
    (untangle B/C passes + delonize) cannot restore Delaunay after these moves:
    its untangle leaves inverted triangles and its Lawson loop then cycles."""
    import ctypes

    from oracle import oracle as O
    L = O.lib()
    pos = np.ascontiguousarray(pos0, np.float32)
    n = len(pos)
    rng = np.random.default_rng(12345)
    out = np.zeros((2 * n, 3), np.int32)
    margin = max(0.02, 6.0 / np.sqrt(n))
    t0 = time.perf_counter()
    steps = 0
    while True:
        x, y = pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64)
        d = cfg["sigma"] * rng.standard_normal((n, 2))
        if cfg["motion"] == 1:
            d[:, 0] += cfg["amp"] * np.sin(2 * np.pi * y)
            d[:, 1] += cfg["amp"] * np.cos(2 * np.pi * x)
            new = (pos + d.astype(np.float32)) % 1.0
        else:
            new = np.clip(pos + d.astype(np.float32), 0.0, 1.0 - 2.0 ** -24)
        pos = np.ascontiguousarray(new, np.float32)
        if periodic:
            k = L.orc_torus_delaunay(pos.ctypes.data_as(ctypes.c_void_p), n, float(margin),
                                     out.ctypes.data_as(ctypes.c_void_p))
            ok = k == 2 * n
        else:
            T = O.HostTri(O.HostTri.EXACT)
            ok = T.add_delaunay_points(pos, 0) == n
        steps += 1
        if time.perf_counter() - t0 > seconds:
            break
    el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "point-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle (restated Host_Triangulation incremental Lawson build, exact predicates) rebuilding "
                      "the exact %s DT of the moved %d points each step: %d steps in %.1f s, 1 thread, host %s, "
                      "last build ok=%s; the reference's untangle+delonize host path does not converge on this "
                      "motion (B/C untangle leaves inverted triangles, Lawson then cycles)"
                      % ("periodic" if periodic else "box", n, steps, el,
                         platform.processor() or platform.machine(), bool(ok))}


def ensemble_sample(counts, k=32):
    """k instances spread over the size range (every len/k-th by size)"""
    order = np.argsort(counts, kind="stable")
    return [int(order[j]) for j in np.linspace(0, len(order) - 1, min(k, len(order))).round().astype(int)]


def plain_ensemble(M, torch, inst_pts, cfg, steps):
    """Per-instance launches on the same GPU for a sample of the ensemble's
    instances: each instance its own triangulation, stepped one after the
    other.  Returns point-steps/s of the plain form (flip-based step with
    FP64-first predicates, variant 1) and of the same fast step as the
    batched run (variant 0) without batching."""
    sample = ensemble_sample(np.array([len(p) for p in inst_pts]))
    dts = []
    for i in sample:
        h = M.HostTriangulation()
        q = inst_pts[i].copy()
        h.addDelaunayPoints(q)
        dts.append((M.DeviceTriangulation(h), len(q)))
    npts = sum(n for _, n in dts)
    out = []
    for variant in (1, 0):
        for d, _ in dts:
            d.step(cfg["motion"], cfg["sigma"], cfg["amp"], seed=cfg["seed"], step=0, variant=variant)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            for d, _ in dts:
                d.step(cfg["motion"], cfg["sigma"], cfg["amp"], seed=cfg["seed"], step=1 + k, variant=variant)
        torch.cuda.synchronize()
        out.append(npts * steps / (time.perf_counter() - t0))
    return out[0], out[1]


def cpu_baseline_ensemble(inst_pos, cfg, seconds):
    """The oracle rebuilding each sampled instance's exact DT after its move,
    instances in turn, one thread."""
    from oracle import oracle as O
    rng = np.random.default_rng(12345)
    pos = [np.ascontiguousarray(p, np.float32) for p in inst_pos]
    t0 = time.perf_counter()
    pts = 0
    rounds = 0
    while time.perf_counter() - t0 < seconds:
        for j, p in enumerate(pos):
            d = cfg["sigma"] * rng.standard_normal(p.shape)
            p = np.ascontiguousarray(np.clip(p + d.astype(np.float32), 0.0, 1.0 - 2.0 ** -24), np.float32)
            T = O.HostTri(O.HostTri.EXACT)
            T.add_delaunay_points(p, 0)
            pos[j] = p
            pts += len(p)
        rounds += 1
    el = time.perf_counter() - t0
    return {"value": pts / el, "unit": "point-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle (restated Host_Triangulation incremental build, exact predicates) rebuilding the "
                      "exact box DT after each move for %d instances spread over the size range (%d points), "
                      "%d rounds in %.1f s, 1 thread, host %s"
                      % (len(pos), sum(len(p) for p in pos), rounds, el, platform.processor() or platform.machine())}


def main():
    a = _args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: every box has one MI355X; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    import torch

    import mcleap_amd as M
    from mcleap_amd.workloads import CONFIGS, ensemble_points, grid_shape, initial_points

    cfg = dict(CONFIGS[a.workload])
    periodic = cfg["domain"] == "torus"
    batch = a.workload == "c5"
    dt = None
    if batch:
        # the ensemble as one mesh: every launch advances all instances
        inst_pts = ensemble_points(a.ninst or None)
        n = sum(len(p) for p in inst_pts)
        t0 = time.time()
        if a.build == "device":
            # every instance's exact triangulation built on the GPU at once (same result and numbering)
            try:
                dt = M.DeviceTriangulation.from_ensemble(inst_pts)
                torch.cuda.synchronize()
                H = dt.h_triangulation
            except M.MclError as e:  # the steps are what is measured: build on the host instead
                print("[bench] device build failed (%s): building on the host" % e, file=sys.stderr, flush=True)
                a.build = "host (device build failed)"
                dt = None
                H = M.HostTriangulationBatch(inst_pts)
        else:
            H = M.HostTriangulationBatch(inst_pts)
        build_s = time.time() - t0
        print("[bench] c5: %d instances, %d points, %s build %.2f s" % (H.ninst, n, a.build, build_s),
              file=sys.stderr, flush=True)
    else:
        xy = initial_points(a.workload, a.n or None)
        n = len(xy)
        t0 = time.time()
        dt = None
        if a.build == "device":
            # the exact triangulation built on the GPU (same result and vertex numbering as the host build)
            try:
                dt = M.DeviceTriangulation.from_points(xy, periodic=periodic)
                torch.cuda.synchronize()
            except M.MclError as e:  # the steps are what is measured: build on the host instead
                print("[bench] device build failed (%s): building on the host" % e, file=sys.stderr, flush=True)
                a.build = "host (device build failed)"
                dt = None
        if dt is None:
            H = M.HostTriangulation(periodic=periodic)
            H.addDelaunayPoints(xy)
        build_s = time.time() - t0
        print("[bench] %s: %d points, %s build %.2f s" % (a.workload, n, a.build, build_s), file=sys.stderr, flush=True)
    if dt is None:
        dt = M.DeviceTriangulation(H)
    nv = dt.m_pos_size
    dev = "cuda"
    ring_off = torch.zeros(nv + 1, dtype=torch.int32, device=dev)
    ring_nbr = torch.zeros(dt.m_he_size, dtype=torch.int32, device=dev)
    radius = cfg["radius"]
    grid = None
    bgrid = None
    if batch:
        # one grid per instance, r_i = 2/sqrt(n_i), sorted and queried together
        radius = "2/sqrt(n_i)"
        rad_i = np.array([2.0 / np.sqrt(max(int(c), 1)) for c in H.counts], np.float32)
        bgrid = M.GridCount2dBatch(H.vstart, np.minimum(rad_i, np.float32(1.0 / 3.0)))
        f_off = torch.zeros(bgrid.numP + 1, dtype=torch.int32, device=dev)
        f_nbr = torch.zeros(bgrid.numP * 32, dtype=torch.int32, device=dev)
    elif radius:
        nc, cs = grid_shape(radius, cfg["domain"])
        flags = M.GRID_MORTON | (M.GRID_PERIODIC if periodic else 0)
        grid = M.GridCount2d(nv, (0.0, 0.0), (cs, cs), (nc, nc), flags)
        f_off = torch.zeros(nv + 1, dtype=torch.int32, device=dev)
        f_nbr = torch.zeros(nv * 32, dtype=torch.int32, device=dev)
    pos_t = dt.pos_tensor()

    def step(k, variant=a.variant):
        st = dt.step(cfg["motion"], cfg["sigma"], cfg["amp"], seed=cfg["seed"], step=k, variant=variant)
        dt.one_ring_csr_async(ring_off, ring_nbr)
        if grid is not None:
            grid.build(pos_t)
            grid.frnn_csr_async(grid.sorted_ptr(), radius, f_off, f_nbr)
        if bgrid is not None:
            bgrid.build(pos_t)
            bgrid.frnn_csr_async(f_off, f_nbr)
        return st

    k = 0
    for _ in range(a.warmup):
        step(k)
        k += 1
    torch.cuda.synchronize()
    dt.kernel_timer(True)
    for g_ in (grid, bgrid):
        if g_ is not None:
            g_.kernel_timer(True)
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        stats.append(step(k))
        k += 1
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = 1e3 * el / a.steps
    print("[bench] timed %d steps: %.3f ms/step" % (a.steps, ms_per_step), file=sys.stderr, flush=True)
    value = n * a.steps / el

    # per-kernel times from the live event timers
    kern = {}
    for name, which in (("lawson", 0), ("check", 1), ("move", 2), ("reloc_apply", 3), ("revert_apply", 4),
                        ("star", 5)):
        ms, cnt = dt.kernel_timer_read(which)
        kern[name] = (ms, cnt)
    kbar = 0.0
    for g_ in (grid, bgrid):
        if g_ is None:
            continue
        for name, which in (("frnn_count", 0), ("frnn_fill", 1), ("grid_build", 2)):
            ms, cnt = g_.kernel_timer_read(which)
            kern[name] = (ms, cnt)
        kbar = float(f_off[-1].item()) / g_.numP
    dt.kernel_timer(False)
    bytes_per_pt = dict(BYTES)
    bytes_per_pt["frnn_fill"] = 4.0 * kbar
    # algorithmic bytes of all timed launches of each kernel class
    algo_total = {k_: bytes_per_pt[k_] * nv * kern[k_][1] for k_ in bytes_per_pt if k_ in kern}
    reloc_units = int(sum(st_["relocated_grid"] for st_ in stats))
    algo_total["reloc_apply"] = RELOC_BYTES * reloc_units
    algo_total["star"] = STAR_BYTES * nv * kern["star"][1]
    dom = max(algo_total, key=lambda k_: kern[k_][0])
    ms_tot, launches = kern[dom]
    avg_ms = ms_tot / max(launches, 1)
    algo = algo_total[dom] / max(launches, 1)
    achieved = algo / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    step_bytes = (212.0 + 4.0 * kbar) * n
    # measured HBM traffic and VALU instructions of the dominant kernel (profiles/)
    traffic, valu_roof = None, None
    try:
        with open(os.path.join(HERE, "profiles", "pmc_traffic.json")) as f:
            pmc = json.load(f).get(a.workload, {}).get(dom)
    except (OSError, ValueError):
        pmc = None
    if pmc and a.n == 0 and a.ninst == 0:
        traffic = pmc["read_bytes"] + pmc["write_bytes"]
        if pmc.get("valu_inst") and avg_ms > 0:
            rate = pmc["valu_inst"] / (avg_ms * 1e-3)
            valu_roof = {"valu_inst_per_launch": pmc["valu_inst"], "achieved": rate, "peak": PEAK_VALU_WAVE_INST,
                         "unit": "wave64 VALU inst/s", "frac": rate / PEAK_VALU_WAVE_INST, "source": pmc["source"]}

    # the result of the last timed step, read through the half-edge arrays
    # (rebuilt here from the stars on the torus): every edge exactly
    # locally Delaunay and no inverted triangle
    final_nd = int(dt.count_non_delaunay())
    final_inv = int(dt.count_inverted())
    if final_nd or final_inv:
        print("[bench] final mesh: %d non-Delaunay edges, %d inverted triangles" % (final_nd, final_inv),
              file=sys.stderr, flush=True)
        sys.exit(1)

    # the plain form (host round trip per round, FP64-first predicates) on the same GPU
    plain = None
    unbatched = None
    if a.plain_steps > 0 and batch:
        plain, unbatched = plain_ensemble(M, torch, inst_pts, cfg, a.plain_steps)
        print("[bench] plain per-instance form: %.3g point-steps/s, unbatched fast form %.3g" % (plain, unbatched),
              file=sys.stderr, flush=True)
    elif a.plain_steps > 0:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(a.plain_steps):
            step(k, variant=1)
            k += 1
        torch.cuda.synchronize()
        plain = n * a.plain_steps / (time.perf_counter() - t1)
        print("[bench] plain form: %.3g point-steps/s" % plain, file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        print("[bench] cpu baseline (%.0f s budget)" % a.cpu_seconds, file=sys.stderr, flush=True)
        p0 = dt.download_positions()
        if batch:
            cpu = cpu_baseline_ensemble([p0[H.vstart[i]:H.vstart[i + 1]] for i in ensemble_sample(H.counts)], cfg,
                                        a.cpu_seconds)
        else:
            cpu = cpu_baseline(p0 if periodic else p0[4:], cfg, periodic, a.cpu_seconds)

    agg = {key: int(sum(s[key] for s in stats)) for key in ("rounds", "flips", "untangle_flips", "reverts",
                                                             "relocated", "relocated_grid", "stalled", "inverted",
                                                             "star_retries", "fallbacks")}
    pin = np.sum([s["pred_incircle"] for s in stats], axis=0)
    por = np.sum([s["pred_orient"] for s in stats], axis=0)
    out = {
        "metric": "point-steps/s restoring exact Delaunay at 1M pts; batched vs plain, same GPU",
        "value": value,
        "unit": "point-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded uniform points, Philox motion on the GPU)",
        "config": {"workload": a.workload, "n_points": n, "domain": cfg["domain"], "sigma": cfg["sigma"],
                   "instances": int(H.ninst) if batch else 1,
                   "amp": cfg["amp"], "frnn_radius": radius, "variant": a.variant,
                   "build": a.build, "build_s": round(build_s, 3),
                   "predicates": "exact: fp32 filter -> fp64 filter -> expansion arithmetic"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "algo_bytes_per_launch": algo, "avg_launch_ms": avg_ms, "launches": launches},
        "valu_roofline": valu_roof,
        "step_roofline": {"algo_bytes_per_step": step_bytes,
                          "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
                          "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / PEAK_HBM_GBS, "kbar": kbar},
        "kernels_ms_per_launch": {k_: (v[0] / max(v[1], 1)) for k_, v in kern.items()},
        "kernels_ms_per_step": {k_: v[0] / a.steps for k_, v in kern.items()},
        "plain_value": plain,
        "unbatched_value": unbatched,
        "fast_vs_plain": (value / plain) if plain else None,
        "per_step_totals": agg,
        "final_mesh": {"non_delaunay_edges": final_nd, "inverted_triangles": final_inv},
        "predicates": {"incircle_stage_counts": pin.tolist(), "orient_stage_counts": por.tolist(),
                       "stages": ["fp32", "fp64", "exact_small", "exact_big"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
