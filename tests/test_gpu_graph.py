"""Whole timesteps replayed from hipGraphs (MCL_OPT_GRAPH) against the same
timesteps launched eagerly: two device triangulations of the same points,
stepped with the same motion, one with MCL_OPT_GRAPH = 1; after every step
the positions, the triangulation, the CSR one-ring and the grid FRNN CSR must
be identical, and the triangulation the exact Delaunay one (oracle).  Graph
replays alternate with eager work on the other object, and both star buffer
parities are replayed several times; a step that gives up on the stars (the
flip-based fallback) must also agree."""
import numpy as np
import pytest

import mcleap_amd as M
from mcleap_amd.workloads import grid_shape
from tests.conftest import assert_at_targets, uniform_points

pytestmark = pytest.mark.gpu


class _Run:
    def __init__(self, H, graph, periodic, rad, give_up=None):
        import torch
        self.dt = M.DeviceTriangulation(H)
        self.dt.set_option(M.OPT_STAR_MIN_VERTICES, 0)
        self.dt.set_option(M.OPT_STAR_ENABLE, 1)
        if give_up:
            self.dt.set_option(M.OPT_STAR_GIVE_UP, give_up)
        self.dt.set_option(M.OPT_GRAPH, int(graph))
        nv = self.dt.m_pos_size
        nc, cs = grid_shape(rad, "torus" if periodic else "box")
        flags = M.GRID_MORTON | (M.GRID_PERIODIC if periodic else 0)
        self.grid = M.GridCount2d(nv, (0.0, 0.0), (cs, cs), (nc, nc), flags)
        self.ring = (torch.zeros(nv + 1, dtype=torch.int32, device="cuda"),
                     torch.zeros(self.dt.m_he_size, dtype=torch.int32, device="cuda"))
        self.frnn = (torch.zeros(nv + 1, dtype=torch.int32, device="cuda"),
                     torch.zeros(nv * 64, dtype=torch.int32, device="cuda"))
        self.rad = rad

    def step(self, motion, sigma, amp, k):
        return self.dt.timestep(motion, sigma, amp=amp, seed=7, step=k, ring=self.ring, grid=self.grid,
                                rad=self.rad, frnn=self.frnn)

    def state(self):
        import torch
        torch.cuda.synchronize()
        pos, v, he, t = self.dt.download()
        return {"pos": pos, "he": he, "t": t, "v": v,
                "ring_off": self.ring[0].cpu().numpy(), "ring_nbr": self.ring[1].cpu().numpy(),
                "frnn_off": self.frnn[0].cpu().numpy(),
                "frnn_nbr": self.frnn[1].cpu().numpy()[: int(self.frnn[0][-1].item())],
                "perm": self.grid.perm().cpu().numpy()}


def _compare(O, a, b, tag, exact_mesh):
    """positions and everything derived from them only (the grid FRNN) are
    always equal; the triangulation is equal as a canonical triangle set, and
    also array by array (half-edge numbering, ring order) while every step so
    far took the star path: the flip-based step's parallel take-backs need not
    number half-edges the same way twice"""
    for key in ("pos", "frnn_off", "frnn_nbr", "perm"):
        assert np.array_equal(a[key], b[key]), "%s: %s differs" % (tag, key)
    assert np.array_equal(O.canonical_triangles(a["he"], a["t"]), O.canonical_triangles(b["he"], b["t"])), tag
    if exact_mesh:
        for key in ("he", "t", "v", "ring_off", "ring_nbr"):
            assert np.array_equal(a[key], b[key]), "%s: %s differs" % (tag, key)


@pytest.mark.parametrize("periodic", [True, False])
def test_graph_timesteps_equal_eager(oracle, periodic):
    O = oracle
    n = 20000 if periodic else 10000
    H = M.HostTriangulation(periodic=periodic)
    H.addDelaunayPoints(uniform_points(n, 44))
    rad = float(np.float32(2.0 / np.sqrt(n)))
    eager, graph = _Run(H, False, periodic, rad), _Run(H, True, periodic, rad)
    motion = (M.MOTION_SHEAR, 2e-4 * np.sqrt(2 ** 20 / n), 1e-3) if periodic else (M.MOTION_GAUSS, 1e-3, 0.0)
    star_steps = 0
    all_star = True
    for k in range(8):
        se = eager.step(motion[0], motion[1], motion[2], k)
        sg = graph.step(motion[0], motion[1], motion[2], k)
        assert se["fallbacks"] == sg["fallbacks"], k
        star = sg["fallbacks"] == 0 and sg["rounds"] == 0
        star_steps += star
        all_star &= star
        a, b = eager.state(), graph.state()
        _compare(O, a, b, "step %d" % k, all_star)
        assert_at_targets(graph.dt, sg, "graph step %d" % k)
        if k % 3 == 2:
            rc, msg = O.check_mesh(b["pos"], b["v"], b["he"], b["t"], periodic=periodic, require_ccw=True)
            assert rc == 0, msg
            assert O.count_non_delaunay(b["pos"], b["he"], periodic=periodic) == 0
    assert star_steps >= (6 if periodic else 4), star_steps


@pytest.mark.parametrize("head", [4, 1])
def test_graph_fallback_equals_eager(oracle, head):
    """a star step that gives up (the first check lists more than 1% of the
    vertices) falls back to the flip-based step in both modes, with equal
    results; the limit is applied on the device by the first retry pass
    (four head passes) or by the host at its first read (one head pass)"""
    n = 20000
    H = M.HostTriangulation(periodic=True)
    H.addDelaunayPoints(uniform_points(n, 45))
    rad = float(np.float32(2.0 / np.sqrt(n)))
    eager, graph = _Run(H, False, True, rad, give_up=1), _Run(H, True, True, rad, give_up=1)
    for r in (eager, graph):
        r.dt.set_option(M.OPT_STAR_HEAD_PASSES, head)
    fallbacks = 0
    for k in range(3):
        se = eager.step(M.MOTION_SHEAR, 1e-3, 1e-3, k)
        sg = graph.step(M.MOTION_SHEAR, 1e-3, 1e-3, k)
        # (after a fallback the star step pauses for a step, in both modes)
        assert se["fallbacks"] == sg["fallbacks"]
        fallbacks += sg["fallbacks"]
        a, b = eager.state(), graph.state()
        # the flip-based step's parallel take-backs may differ in their
        # intermediate meshes; the result is the unique Delaunay triangulation
        assert np.array_equal(a["pos"], b["pos"])
        assert np.array_equal(oracle.canonical_triangles(a["he"], a["t"]), oracle.canonical_triangles(b["he"], b["t"]))
        assert_at_targets(graph.dt, sg, "fallback step %d" % k)
    assert fallbacks >= 1


def test_tuning_table_applied_at_create(tmp_path):
    """an entry of the autotuner's table sets the options of a triangulation
    of its class and size at creation (and grid_block of a grid), and leaves
    other classes at the defaults"""
    import json
    import os
    import subprocess
    import sys
    table = {"entries": [{"domain": "torus", "log2_nv": 14, "graph": 1, "star_head_passes": 3, "grid_block": 128},
                         {"domain": "box", "log2_nv": 14, "tail_threshold": 4096, "reloc_tail": 256}]}
    path = tmp_path / "t.json"
    path.write_text(json.dumps(table))
    code = r"""
import json, numpy as np, mcleap_amd as M
from tests.conftest import uniform_points
out = {}
for periodic in (True, False):
    H = M.HostTriangulation(periodic=periodic)
    H.addDelaunayPoints(uniform_points(16000, 3))
    dt = M.DeviceTriangulation(H)
    out["torus" if periodic else "box"] = dt.options()
g = M.GridCount2d(16000, (0.0, 0.0), (0.01, 0.01), (100, 100), M.GRID_MORTON | M.GRID_PERIODIC)
out["grid_block_torus"] = g.block()
g2 = M.GridCount2d(16000, (0.0, 0.0), (0.01, 0.01), (100, 100), M.GRID_MORTON)
out["grid_block_box"] = g2.block()
print(json.dumps(out))
"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=root,
                       env=dict(os.environ, MCL_TUNING=str(path)), timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["torus"]["graph"] == 1 and out["torus"]["star_head_passes"] == 3
    assert out["torus"]["tail_threshold"] == 8192
    assert out["box"]["tail_threshold"] == 4096 and out["box"]["reloc_tail"] == 256 and out["box"]["graph"] == 0
    assert out["grid_block_torus"] == 128 and out["grid_block_box"] == 256


@pytest.mark.parametrize("periodic,head,graph", [(True, 1, False), (True, 4, False), (True, 1, True), (False, 1, False)])
def test_timestep_tail_equals_fresh_queries(oracle, periodic, head, graph):
    """mcl_tri_timestep runs the grid FRNN part of its tail on a second stream
    from the star step's first check on (with graphs: on the caller's stream
    at the counter read), before the stars have settled, and enqueues the
    one-ring at the counter read (each again after the step when the
    positions or the stars changed later).  After every
    step the ring and FRNN arrays it wrote must equal queries made afresh on
    the resulting mesh and positions.  One head pass makes most steps settle
    only after that read (the path that enqueues the one-ring late and may
    repeat the FRNN); the box case (moves of four mean spacings) clamps
    points onto the sides and corners of the unit square, where exact
    duplicates send vertices back, and its first step defers relocations
    whose star passed 64 triangles in the relocation phase's mesh (relocated
    again after the Lawson rounds: no vertex may stall)."""
    import torch
    n = 20000
    H = M.HostTriangulation(periodic=periodic)
    H.addDelaunayPoints(uniform_points(n, 46))
    rad = float(np.float32(2.0 / np.sqrt(n)))
    r = _Run(H, graph, periodic, rad)
    r.dt.set_option(M.OPT_STAR_HEAD_PASSES, head)
    nv = r.dt.m_pos_size
    nc, cs = grid_shape(rad, "torus" if periodic else "box")
    fresh = M.GridCount2d(nv, (0.0, 0.0), (cs, cs), (nc, nc), M.GRID_MORTON | (M.GRID_PERIODIC if periodic else 0))
    off2 = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    nbr2 = torch.zeros(nv * 64, dtype=torch.int32, device="cuda")
    roff2 = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    rnbr2 = torch.zeros(r.dt.m_he_size, dtype=torch.int32, device="cuda")
    motion = (M.MOTION_SHEAR, 2e-4 * np.sqrt(2 ** 20 / n), 1e-3) if periodic else (M.MOTION_GAUSS, 3e-2, 0.0)
    late = 0
    for k in range(6):
        st = r.step(motion[0], motion[1], motion[2], k)
        late += st["iterations"] > head
        torch.cuda.synchronize()
        fresh.build(r.dt.pos_tensor())
        tot = fresh.frnn_csr(fresh.sorted_ptr(), rad, off2, nbr2)
        assert torch.equal(r.frnn[0], off2), "step %d: FRNN offsets differ" % k
        assert torch.equal(r.frnn[1][:tot], nbr2[:tot]), "step %d: FRNN lists differ" % k
        tr = r.dt.one_ring_csr(roff2, rnbr2)
        assert torch.equal(r.ring[0], roff2), "step %d: ring offsets differ" % k
        assert torch.equal(r.ring[1][:tr], rnbr2[:tr]), "step %d: rings differ" % k
        assert_at_targets(r.dt, st, "step %d" % k)
    pos, v, he, t = r.dt.download()
    assert O_check(oracle, pos, v, he, t, periodic)
    if head == 1 and periodic:
        assert late >= 1, "no step settled after the first read"


def O_check(O, pos, v, he, t, periodic):
    rc, msg = O.check_mesh(pos, v, he, t, periodic=periodic, require_ccw=True)
    assert rc == 0, msg
    assert O.count_non_delaunay(pos, he, periodic=periodic) == 0
    return True
