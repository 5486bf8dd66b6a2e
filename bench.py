"""Kinetic Delaunay benchmark on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload c2] [--gpus 1]

A step is one pass of the hot path over the whole point set: move every
point, restore the exact Delaunay triangulation (on the torus: per-vertex
star recomputation with consistency checks; in the box: repair flips /
take-backs / relocation / Lawson flips), one-ring CSR, and the uniform-grid
FRNN rebuild + query (CSR), issued as one mcl_tri_timestep call, with the
library's launch options (MCL_OPT_GRAPH from its tuning table, eager
launches by default; --graph 1 replays the star step's head and the tail
from hipGraphs).  Inputs live in HBM before the timed region; the K timed steps are
bracketed by torch.cuda.synchronize().  Then --timer-steps further steps run
with eager launches and per-kernel HIP event timers, for the roofline.  After the timed
steps the final mesh is read through the half-edge arrays and checked (every
edge exactly locally Delaunay, no inverted triangle, every vertex at its
target except exact-duplicate targets); a failed check exits non-zero.

The default run measures the 1M-point torus workload (c2) as the headline and
then the 4096-instance ensemble (c5) as the "ensemble" sub-record: batched
point-steps/s against the plain per-instance form on the same GPU.

roofline: the dominant kernel class (largest GPU time) among the library's
per-kernel HIP event timers on the caller's stream (the grid sort and FRNN run
on the library's second stream beside the step on large meshes, sharing the
GPU; relocation and take-back have no phase in the model; the largest class
overall is reported as largest_class), credited with the algorithmic bytes of the phase
it performs, per point: move 16 B, inversion check 32 B, Delaunay check 56 B,
the first star pass (which replaces both checks) 88 B, grid build 43 B, FRNN
count 13 B, FRNN fill 4*k B (k = mean neighbours).  Relocation and take-back
launches have no such phase and are credited nothing.  gather_model: the
star pass's bytes if every gathered row counted once per reader (390 B per
point), reported apart.  traffic: the HBM bytes per launch of that kernel
measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a copy
kernel), from profiles/pmc_traffic.json.  valu_roofline: the same launch's
VALU instruction count (SQ_INSTS_VALU, same file) over the measured launch
time against the issue peak (256 CUs x 4 SIMDs x one wave64 instruction per 4
cycles at 2.4 GHz).  step_roofline: the whole step credited with
(212 + 4k) B per point.
predicate_check: every interior edge of the final mesh (the incircle calls of
a Lawson round over all edges) and every triangle (the orientation calls of
the inversion check), evaluated by the step's cascade, the plain path's
FP64-first cascade, a plain FP64 determinant and exact expansion arithmetic
alone: sign disagreements with the exact signs, and the cascade's stages.
cpu_baseline: the oracle's exact host triangulation rebuilt after each move,
one thread, with the host's CPU model and core count.
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
PEAK_VALU_WAVE_INST = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second
# algorithmic bytes per point of each timed kernel class (the per-phase model)
PHASE_BYTES = {"move": 16, "check": 32, "lawson": 56, "star": 88, "grid_build": 43, "frnn_count": 13}
STAR_GATHER_BYTES = 390  # per vertex: own star 28 + 6 neighbour stars 6*28 + ~19 candidate positions 152 + new star 28 + pos 8


def host_cpu():
    """CPU model name and online core count of this host"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 0


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
    ap.add_argument("--no-ensemble", action="store_true", help="skip the c5 sub-record of the default c2 run")
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--no-pred-check", action="store_true")
    ap.add_argument("--graph", type=int, default=-1, choices=(-1, 0, 1),
                    help="replay each timestep's star head and tail from hipGraphs (MCL_OPT_GRAPH; "
                         "-1: the library's setting)")
    ap.add_argument("--timer-steps", type=int, default=10,
                    help="eager steps timed after the run with per-kernel event timers (the roofline figures)")
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


def _cpu_record(value, sample):
    model, ncpu = host_cpu()
    return {"value": value, "unit": "point-steps/s", "cores": 1, "kind": "port", "cpu_model": model, "nproc": ncpu,
            "sample": sample + "; host CPU %s (%d cores online), 1 thread used" % (model, ncpu)}


def _keep_distinct(old, new):
    """the moved points with exact coincidences resolved as the GPU step does:
    of the points on one spot the smallest index keeps it, the others stay
    at their previous positions (repeated while that makes new coincidences)"""
    new = new.copy()
    for _ in range(8):
        _, first, inv = np.unique(new.view(np.uint64).ravel(), return_index=True, return_inverse=True)
        back = first[inv.ravel()] != np.arange(len(new))
        if not back.any():
            break
        new[back] = old[back]
    return new


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
    ok = True
    while True:
        x, y = pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64)
        d = cfg["sigma"] * rng.standard_normal((n, 2))
        if cfg["motion"] == 1:
            d[:, 0] += cfg["amp"] * np.sin(2 * np.pi * y)
            d[:, 1] += cfg["amp"] * np.cos(2 * np.pi * x)
            new = (pos + d.astype(np.float32)) % 1.0
        else:
            new = np.clip(pos + d.astype(np.float32), 0.0, 1.0 - 2.0 ** -24)
        pos = _keep_distinct(pos, np.ascontiguousarray(new, np.float32))
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
    return _cpu_record(n * steps / el,
                       "oracle (restated Host_Triangulation incremental Lawson build, exact predicates) rebuilding "
                       "the exact %s DT of the moved %d points from scratch each step: %d steps in %.1f s, last "
                       "build ok=%s; the reference's untangle+delonize host path does not converge on this motion "
                       "(B/C untangle leaves inverted triangles, Lawson then cycles)"
                       % ("periodic" if periodic else "box", n, steps, el, bool(ok)))


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
    return out[0], out[1], len(sample), npts


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
    return _cpu_record(pts / el,
                       "oracle (restated Host_Triangulation incremental build, exact predicates) rebuilding the "
                       "exact box DT from scratch after each move for %d instances spread over the size range (%d "
                       "points), %d rounds in %.1f s" % (len(pos), sum(len(p) for p in pos), rounds, el))


def predicate_check(M, torch, dt):
    """Sign disagreements with exact arithmetic on the final mesh's predicate
    calls: incircle of every interior edge (inCircle(org(e), op(twin),
    org(twin), op(e)), kernels.cuh:287-292) and orientation of every triangle,
    by the step's FP32-first cascade, the plain path's FP64-first cascade and
    a plain FP64 determinant, against exact expansions alone."""
    pos, v, he, t = dt.download()
    P = pos[v[:, 0]]
    h = np.arange(0, len(he), 2)
    inner = (he[h, 2] >= 0) & (he[h + 1, 2] >= 0)
    h = h[inner]
    quad = np.stack([P[he[h, 1]], P[he[h + 1, 3]], P[he[h + 1, 1]], P[he[h, 3]]], 1)
    h0 = t
    h1 = he[h0, 0]
    h2 = he[h1, 0]
    tri = np.stack([P[he[h0, 1]], P[he[h1, 1]], P[he[h2, 1]]], 1)
    per = int(dt.periodic)
    L = M.lib()

    def ev(kind, arr):
        n = len(arr)
        d = torch.from_numpy(np.ascontiguousarray(arr, np.float32).reshape(-1, 2)).cuda()
        sg = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        st = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        rc = L.mcl_eval_predicates(kind, n, d.data_ptr(), per, sg.data_ptr(), st.data_ptr(), None)
        if rc != 0:
            raise M.MclError(rc, "mcl_eval_predicates")
        return sg.cpu().numpy()[:n], st.cpu().numpy()[:n]

    ex_i, _ = ev(6, quad)
    ex_o, _ = ev(7, tri)
    out = {"incircle_cases": int(len(quad)), "orient_cases": int(len(tri))}
    for name, kind in (("fp32_first_cascade", 1), ("fp64_first_cascade", 3), ("plain_fp64", 5)):
        s_, st = ev(kind, quad)
        rec = {"disagreements": int((s_ != ex_i).sum())}
        if kind != 5:
            rec["stage_fractions"] = (np.bincount(st, minlength=4)[:4] / max(len(st), 1)).tolist()
        out["incircle_" + name] = rec
    for name, kind in (("fp32_first_cascade", 0), ("fp64_first_cascade", 2), ("plain_fp64", 4)):
        s_, st = ev(kind, tri)
        rec = {"disagreements": int((s_ != ex_o).sum())}
        if kind != 4:
            rec["stage_fractions"] = (np.bincount(st, minlength=4)[:4] / max(len(st), 1)).tolist()
        out["orient_" + name] = rec
    out["stages"] = ["fp32", "fp64", "exact_small", "exact_big"]
    return out


def fail(msg):
    print("[bench] FAILED: " + msg, file=sys.stderr, flush=True)
    sys.exit(1)


def run_workload(a, M, torch, workload, steps, warmup, plain_steps, cpu_seconds, pred_check):
    """one workload: build, warm up, time `steps` steps, check, compare"""
    from mcleap_amd import tuning
    from mcleap_amd.workloads import CONFIGS, ensemble_points, grid_shape, initial_points

    cfg = dict(CONFIGS[workload])
    periodic = cfg["domain"] == "torus"
    batch = workload == "c5"
    dt = None
    H = None
    inst_pts = None
    t0 = time.time()
    if batch:
        # the ensemble as one mesh: every launch advances all instances
        inst_pts = ensemble_points(a.ninst or None)
        n = sum(len(p) for p in inst_pts)
        if a.build == "device":
            try:
                dt = M.DeviceTriangulation.from_ensemble(inst_pts)
            except M.MclError as e:
                fail("device build of the ensemble: %s (--build host builds it on the host)" % e)
            torch.cuda.synchronize()
            H = dt.h_triangulation
        else:
            H = M.HostTriangulationBatch(inst_pts)
    else:
        xy = initial_points(workload, a.n or None)
        n = len(xy)
        if a.build == "device":
            try:
                dt = M.DeviceTriangulation.from_points(xy, periodic=periodic)
            except M.MclError as e:
                fail("device build of %s: %s (--build host builds it on the host)" % (workload, e))
            torch.cuda.synchronize()
        else:
            H = M.HostTriangulation(periodic=periodic)
            H.addDelaunayPoints(xy)
    if dt is None:
        dt = M.DeviceTriangulation(H)
    build_s = time.time() - t0
    print("[bench] %s: %d points, %s build %.2f s" % (workload, n, a.build, build_s), file=sys.stderr, flush=True)
    nv = dt.m_pos_size
    dev = "cuda"
    ring_off = torch.zeros(nv + 1, dtype=torch.int32, device=dev)
    ring_nbr = torch.zeros(dt.m_he_size, dtype=torch.int32, device=dev)
    radius = cfg["radius"]
    grid = bgrid = None
    cells = 1
    pos_t = dt.pos_tensor()
    if batch:
        # one grid per instance, r_i = 2/sqrt(n_i), sorted and queried together
        radius = "2/sqrt(n_i)"
        rad_i = np.array([2.0 / np.sqrt(max(int(c), 1)) for c in H.counts], np.float32)
        bgrid = M.GridCount2dBatch(H.vstart, np.minimum(rad_i, np.float32(1.0 / 3.0)))
        f_off = torch.zeros(bgrid.numP + 1, dtype=torch.int32, device=dev)
        bgrid.build(pos_t)
        total = bgrid.frnn_csr(f_off)
    elif radius:
        # cells of the radius, or of a fraction of it when the tuning table says so
        cells = tuning.cells_per_radius("torus" if periodic else "box", nv)
        nc, cs = grid_shape(radius / cells, cfg["domain"])
        flags = M.GRID_MORTON | (M.GRID_PERIODIC if periodic else 0)
        grid = M.GridCount2d(nv, (0.0, 0.0), (cs, cs), (nc, nc), flags)
        f_off = torch.zeros(nv + 1, dtype=torch.int32, device=dev)
        grid.build(pos_t)
        total = grid.frnn_csr(grid.sorted_ptr(), radius, f_off)
    if grid is not None or bgrid is not None:
        # the neighbour buffer sized from a count pass (25% + 1024 of head room
        # for the motion); the run fails if a step ever needs more
        f_cap = int(total * 1.25) + 1024
        f_nbr = torch.zeros(f_cap, dtype=torch.int32, device=dev)

    def step(k, variant=a.variant):
        # the whole timestep in one call: kinetic step, CSR one-ring, grid sort + FRNN
        if bgrid is None:
            return dt.timestep(cfg["motion"], cfg["sigma"], cfg["amp"], seed=cfg["seed"], step=k, variant=variant,
                               ring=(ring_off, ring_nbr), grid=grid, rad=radius or 0.0,
                               frnn=(f_off, f_nbr) if grid is not None else None)
        return dt.timestep(cfg["motion"], cfg["sigma"], cfg["amp"], seed=cfg["seed"], step=k, variant=variant,
                           ring=(ring_off, ring_nbr), bgrid=bgrid, frnn=(f_off, f_nbr))

    def timers(on):
        dt.kernel_timer(on)
        for g_ in (grid, bgrid):
            if g_ is not None:
                g_.kernel_timer(on)

    if a.graph >= 0:
        dt.set_option(M.OPT_GRAPH, a.graph)
    graph = bool(dt.get_option(M.OPT_GRAPH))
    k = 0
    for _ in range(warmup):
        step(k)
        k += 1
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(step(k))
        k += 1
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = 1e3 * el / steps
    value = n * steps / el
    print("[bench] %s timed %d steps: %.3f ms/step%s" % (workload, steps, ms_per_step, " (graphs)" if graph else ""),
          file=sys.stderr, flush=True)
    # per-kernel GPU times: further steps launched eagerly with event timers
    dt.set_option(M.OPT_GRAPH, 0)
    timers(True)
    tsteps = max(1, a.timer_steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(tsteps):
        stats.append(step(k))
        k += 1
    torch.cuda.synchronize()
    eager_ms = 1e3 * (time.perf_counter() - t1) / tsteps
    print("[bench] %s %d eager steps with kernel timers: %.3f ms/step" % (workload, tsteps, eager_ms),
          file=sys.stderr, flush=True)

    # per-kernel times from the live event timers
    kern = {}
    for name, which in (("lawson", 0), ("check", 1), ("move", 2), ("reloc_apply", 3), ("revert_apply", 4),
                        ("star", 5)):
        kern[name] = dt.kernel_timer_read(which)
    kern = {k_: v_ for k_, v_ in kern.items() if v_[1] > 0}
    kbar = 0.0
    for g_ in (grid, bgrid):
        if g_ is None:
            continue
        for name, which in (("frnn_count", 0), ("frnn_fill", 1), ("grid_build", 2)):
            kern[name] = g_.kernel_timer_read(which)
        last_pairs = int(f_off[-1].item())
        kbar = last_pairs / g_.numP
        if last_pairs > f_cap:
            fail("FRNN output needs %d entries, the buffer holds %d" % (last_pairs, f_cap))
    timers(False)
    bytes_per_pt = dict(PHASE_BYTES)
    bytes_per_pt["frnn_fill"] = 4.0 * kbar
    # algorithmic bytes of all timed launches of each kernel class (none for
    # relocation and take-back, which no phase of the model performs)
    algo_total = {k_: bytes_per_pt.get(k_, 0.0) * nv * kern[k_][1] for k_ in kern}
    # the dominant kernel: the largest GPU time among the phases of the model
    # on the caller's stream.  The grid sort and FRNN launches run on the
    # library's second stream beside the step (meshes of star_min_vertices or
    # more, eager launches), where their event times include sharing the GPU;
    # relocation and take-back have no algorithmic bytes.  The largest class
    # left out is reported beside it.
    side = (not graph) and (grid is not None or bgrid is not None) and \
        nv >= dt.get_option(M.OPT_STAR_MIN_VERTICES)
    side_names = ("grid_build", "frnn_count", "frnn_fill") if side else ()
    modelled = [k_ for k_ in kern if k_ not in side_names and bytes_per_pt.get(k_, 0.0) > 0]
    dom = max(modelled or list(kern), key=lambda k_: kern[k_][0])
    largest = max(kern, key=lambda k_: kern[k_][0])
    ms_tot, launches = kern[dom]
    avg_ms = ms_tot / max(launches, 1)
    algo = algo_total[dom] / max(launches, 1)
    achieved = algo / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    step_bytes = (212.0 + 4.0 * kbar) * n
    traffic, valu_roof = None, None
    try:
        with open(os.path.join(HERE, "profiles", "pmc_traffic.json")) as f:
            pmc = json.load(f).get(workload, {}).get(dom)
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
    # locally Delaunay, no inverted triangle, every vertex at its target
    final_nd = int(dt.count_non_delaunay())
    final_inv = int(dt.count_inverted())
    pos_end = dt.download_positions()
    tgt_end = dt.download_targets()
    off_target = int(np.any(pos_end != tgt_end, axis=1).sum())
    agg = {key: int(sum(s_[key] for s_ in stats)) for key in ("rounds", "flips", "untangle_flips", "reverts",
                                                               "relocated", "relocated_grid", "stalled",
                                                               "duplicates", "inverted", "star_retries", "fallbacks")}
    if final_nd or final_inv or agg["stalled"] or off_target != stats[-1]["duplicates"]:
        fail("%s final mesh: %d non-Delaunay edges, %d inverted triangles, %d stalled, %d off target"
             % (workload, final_nd, final_inv, agg["stalled"], off_target))
    preds = None
    if pred_check:
        preds = predicate_check(M, torch, dt)
        bad = preds["incircle_fp32_first_cascade"]["disagreements"] + preds["orient_fp32_first_cascade"]["disagreements"]
        print("[bench] %s predicate check: %d incircle / %d orient cases, cascade disagreements %d, plain FP64 %d"
              % (workload, preds["incircle_cases"], preds["orient_cases"], bad,
                 preds["incircle_plain_fp64"]["disagreements"] + preds["orient_plain_fp64"]["disagreements"]),
              file=sys.stderr, flush=True)
        if bad:
            fail("the step's predicate cascade disagrees with exact arithmetic")

    # the plain form (host round trip per round, FP64-first predicates) on the same GPU
    plain = unbatched = None
    plain_sample = None
    if plain_steps > 0 and batch:
        plain, unbatched, ninst_s, npts_s = plain_ensemble(M, torch, inst_pts, cfg, plain_steps)
        plain_sample = "%d instances spread over the size range (%d points), %d steps each" % (ninst_s, npts_s,
                                                                                             plain_steps)
        print("[bench] plain per-instance form: %.3g point-steps/s, unbatched fast form %.3g" % (plain, unbatched),
              file=sys.stderr, flush=True)
    elif plain_steps > 0:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(plain_steps):
            step(k, variant=1)
            k += 1
        torch.cuda.synchronize()
        plain = n * plain_steps / (time.perf_counter() - t1)
        plain_sample = "the same mesh, %d steps" % plain_steps
        print("[bench] plain form: %.3g point-steps/s" % plain, file=sys.stderr, flush=True)

    cpu = None
    if cpu_seconds > 0:
        print("[bench] cpu baseline (%.0f s budget)" % cpu_seconds, file=sys.stderr, flush=True)
        p0 = dt.download_positions()
        if batch:
            cpu = cpu_baseline_ensemble([p0[H.vstart[i]:H.vstart[i + 1]] for i in ensemble_sample(H.counts)], cfg,
                                        cpu_seconds)
        else:
            cpu = cpu_baseline(p0 if periodic else p0[4:], cfg, periodic, cpu_seconds)

    pin = np.sum([s_["pred_incircle"] for s_ in stats], axis=0)
    por = np.sum([s_["pred_orient"] for s_ in stats], axis=0)
    hbm_frac = achieved / PEAK_HBM_GBS
    rec = {
        "value": value,
        "ms_per_step": ms_per_step,
        "config": {"workload": workload, "n_points": n, "domain": cfg["domain"], "sigma": cfg["sigma"],
                   "instances": int(H.ninst) if batch else 1,
                   "amp": cfg["amp"], "frnn_radius": radius, "variant": a.variant,
                   "build": a.build, "build_s": round(build_s, 3),
                   "predicates": "exact: fp32 filter -> fp64 filter -> expansion arithmetic",
                   "tuning": {"table": os.path.relpath(tuning.table_path(), HERE) if tuning.table_path() else None,
                              "options": dt.options(),
                              "grid_block": grid.block() if grid is not None else None,
                              "frnn_cells_per_radius": cells if grid is not None else None},
                   "half_edges": "star path: rebuilt from the stars only when read (not inside the timed steps)"
                   if periodic else "maintained by every step"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": hbm_frac, "traffic": traffic, "algo_bytes_per_launch": algo,
                     "algo_bytes_per_point": bytes_per_pt.get(dom, 0.0), "avg_launch_ms": avg_ms,
                     "launches": launches,
                     "binding": "valu" if valu_roof and valu_roof["frac"] > hbm_frac else "hbm",
                     "largest_class": largest, "side_stream": list(side_names)},
        "valu_roofline": valu_roof,
        "step_roofline": {"algo_bytes_per_step": step_bytes,
                          "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
                          "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / PEAK_HBM_GBS, "kbar": kbar},
        "kernels_ms_per_launch": {k_: (v_[0] / max(v_[1], 1)) for k_, v_ in kern.items()},
        "kernels_ms_per_step": {k_: v_[0] / tsteps for k_, v_ in kern.items()},
        "graph": graph,
        "eager_ms_per_step": eager_ms,
        "plain_value": plain,
        "plain_sample": plain_sample,
        "unbatched_value": unbatched,
        "fast_vs_plain": (value / plain) if plain else None,
        "per_step_totals": agg,
        "final_mesh": {"non_delaunay_edges": final_nd, "inverted_triangles": final_inv,
                       "off_target_vertices": off_target, "duplicate_targets": stats[-1]["duplicates"]},
        "frnn": {"pairs_last_step": int(f_off[-1].item()) if (grid is not None or bgrid is not None) else 0,
                 "capacity": f_cap if (grid is not None or bgrid is not None) else 0},
        "predicates": {"incircle_stage_counts": pin.tolist(), "orient_stage_counts": por.tolist(),
                       "stages": ["fp32", "fp64", "exact_small", "exact_big"]},
        "predicate_check": preds,
        "cpu_baseline": cpu,
    }
    if dom == "star":
        rec["roofline"]["gather_model"] = {"bytes_per_point": STAR_GATHER_BYTES,
                                           "achieved": STAR_GATHER_BYTES * nv / (avg_ms * 1e-3) / 1e9}
    return rec


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

    cpu_s = 0.0 if a.no_cpu else a.cpu_seconds
    rec = run_workload(a, M, torch, a.workload, a.steps, a.warmup, a.plain_steps, cpu_s, not a.no_pred_check)
    ens = None
    if a.workload == "c2" and not a.no_ensemble and a.n == 0:
        gc.collect()
        torch.cuda.empty_cache()
        e = run_workload(a, M, torch, "c5", a.ensemble_steps, 2, 3, 0.0 if a.no_cpu else min(10.0, a.cpu_seconds),
                         False)
        ens = {"workload": "c5", "metric": "point-steps/s, 4096 instances advanced by shared launches",
               "value": e["value"], "ms_per_step": e["ms_per_step"], "steps": a.ensemble_steps,
               "n_points": e["config"]["n_points"], "instances": e["config"]["instances"],
               "build_s": e["config"]["build_s"], "plain_value": e["plain_value"], "plain_sample": e["plain_sample"],
               "unbatched_value": e["unbatched_value"], "batched_vs_plain": e["fast_vs_plain"],
               "batched_vs_unbatched": (e["value"] / e["unbatched_value"]) if e["unbatched_value"] else None,
               "roofline": e["roofline"], "kernels_ms_per_step": e["kernels_ms_per_step"],
               "per_step_totals": e["per_step_totals"], "final_mesh": e["final_mesh"], "cpu_baseline": e["cpu_baseline"]}
    out = {
        "metric": "point-steps/s restoring exact Delaunay at 1M pts; batched vs plain, same GPU",
        "value": rec["value"],
        "unit": "point-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded uniform points, Philox motion on the GPU)",
    }
    for key in ("config", "graph", "eager_ms_per_step", "roofline", "valu_roofline", "step_roofline",
                "kernels_ms_per_launch", "kernels_ms_per_step",
                "plain_value", "plain_sample", "fast_vs_plain", "unbatched_value", "per_step_totals", "final_mesh",
                "frnn", "predicates", "predicate_check", "cpu_baseline"):
        out[key] = rec[key]
    out["ensemble"] = ens
    print(json.dumps(out))


if __name__ == "__main__":
    main()
