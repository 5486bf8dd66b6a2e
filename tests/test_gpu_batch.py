"""GPU parity of the batched ensemble: many independent box triangulations
held as one mesh (HostTriangulationBatch) and advanced by the same launches.

After every step the downloaded union mesh is split back into its instances
(half-edges by origin vertex, renumbered per instance).  Each instance must
(1) satisfy the half-edge invariants and Euler's formula on its own, with
every triangle strictly CCW and no half-edge, triangle or vertex reference
into another instance, (2) have no edge failing the exact incircle test, and
(3) equal, as a canonical triangle set, the oracle's from-scratch exact
Delaunay triangulation of that instance's downloaded positions.
"""
import numpy as np
import pytest

import mcleap_amd as M
from tests.conftest import assert_at_targets
from mcleap_amd.workloads import ensemble_points
from tests.test_host_batch import batch_sizes

pytestmark = pytest.mark.gpu


def _submesh(B, pos, v, he, t, i):
    """instance i of the union as its own AoS mesh (vertex ids: corners 0..3,
    then the points in the instance's vertex order)"""
    verts = B.instance_vertices(i)
    vloc = np.full(len(pos), -1, np.int64)
    vloc[verts] = np.arange(len(verts))
    hsel = np.nonzero(vloc[he[:, 1]] >= 0)[0]
    hloc = np.full(len(he), -1, np.int64)
    hloc[hsel] = np.arange(len(hsel))
    tids = np.unique(he[hsel, 2][he[hsel, 2] >= 0])
    tloc = np.full(len(t), -1, np.int64)
    tloc[tids] = np.arange(len(tids))
    sub = he[hsel]
    nxt = hloc[sub[:, 0]]
    tt = np.where(sub[:, 2] >= 0, tloc[np.maximum(sub[:, 2], 0)], sub[:, 2])
    op = np.where(sub[:, 3] >= 0, vloc[np.maximum(sub[:, 3], 0)], sub[:, 3])
    assert np.all(nxt >= 0) and np.all(tt >= -1) and np.all((sub[:, 3] < 0) | (op >= 0)), "cross-instance reference"
    # twins stay pairs (2e, 2e+1) after the selection
    assert np.all(hsel[0::2] % 2 == 0) and np.all(hsel[1::2] == hsel[0::2] + 1), "twin pairs split"
    he_l = np.stack([nxt, vloc[sub[:, 1]], tt, op], 1).astype(np.int32)
    t_l = hloc[t[tids]]
    assert np.all(t_l >= 0)
    v_he = hloc[v[verts, 1]]
    assert np.all(v_he >= 0), "vertex half-edge in another instance"
    v_l = np.stack([np.arange(len(verts)), v_he], 1).astype(np.int32)
    return pos[verts], v_l, he_l, t_l.astype(np.int32)


def _check_batch(O, B, dt, tag, instances=None):
    pos, v, he, t = dt.download()
    for i in (range(B.ninst) if instances is None else instances):
        p_i, v_i, h_i, t_i = _submesh(B, pos, v, he, t, i)
        rc, msg = O.check_mesh(p_i, v_i, h_i, t_i, periodic=False, require_ccw=True)
        assert rc == 0, "%s instance %d: %s" % (tag, i, msg)
        if len(p_i) == 4:
            # no points: the square's two triangles; its 4 corners are exactly
            # cocircular and either diagonal may stand
            continue
        assert O.count_non_delaunay(p_i, h_i) == 0, "%s instance %d" % (tag, i)
        ref = O.delaunay(p_i[4:])
        p2, v2, h2, t2 = ref.export()
        assert np.array_equal(O.canonical_triangles(h_i, t_i), O.canonical_triangles(h2, t2)), \
            "%s instance %d (n=%d)" % (tag, i, len(p_i) - 4)
    return pos


@pytest.mark.parametrize("star", [True, False])
def test_batch_steps(oracle, star):
    """24 instances (n = 0 .. 3000) advanced together: star recomputation
    (corner fans of every instance, an empty instance's fans kept) or the
    flip-based step, checked per instance after every step."""
    O = oracle
    pts = ensemble_points(seed=11, sizes=batch_sizes())
    B = M.HostTriangulationBatch(pts)
    dt = M.DeviceTriangulation(B)
    dt.set_option(M.OPT_STAR_MIN_VERTICES, 0)
    dt.set_option(M.OPT_STAR_ENABLE, 1 if star else 0)
    prev = dt.download_positions()
    nc = 4 * B.ninst
    star_steps = 0
    nsteps = 8
    for k in range(nsteps):
        st = dt.step(M.MOTION_GAUSS, 2e-3, seed=5, step=k)
        print("step", k, st, flush=True)
        pos = _check_batch(O, B, dt, "star %s step %d" % (star, k))
        assert np.array_equal(pos[:nc], prev[:nc])  # corners never move
        assert np.all(pos[nc:] >= 0) and np.all(pos[nc:] < 1)
        assert np.any(pos[nc:] != prev[nc:], axis=1).mean() > 0.99
        assert_at_targets(dt, st, "star %s step %d" % (star, k))
        star_steps += star and st["fallbacks"] == 0 and st["rounds"] == 0
        prev = pos
    if star:
        assert star_steps >= nsteps // 2, star_steps


def test_batch_c5_sizes(oracle):
    """64 instances with the C5 size law (n = 256 .. 65536), sigma = 5e-4:
    exact after every step, whichever path each step took"""
    pts = ensemble_points(64, seed=5)
    B = M.HostTriangulationBatch(pts)
    dt = M.DeviceTriangulation(B)
    dt.set_option(M.OPT_STAR_MIN_VERTICES, 0)
    for k in range(6):
        st = dt.step(M.MOTION_GAUSS, 5e-4, seed=5, step=k)
        print("step", k, st, flush=True)
        assert dt.count_inverted() == 0 and dt.count_non_delaunay() == 0
        assert_at_targets(dt, st, "step %d" % k)
    _check_batch(O=oracle, B=B, dt=dt, tag="after 6 steps", instances=[0, 1, 17, 63])


@pytest.mark.parametrize("star,sigma", [(True, 2e-3), (False, 2e-3), (False, 3e-2)])
def test_batch_timestep_tail_equals_fresh_queries(oracle, star, sigma):
    """mcl_tri_timestep_batch: the ensemble's step with its CSR one-ring and
    the batched grids' FRNN query in one call (the FRNN part runs on a
    second stream on the targets, beside the step, and again after it when a
    vertex is not at its target: moves of sigma = 3e-2 clamp points onto the
    sides and corners of the unit square, where exact duplicates stay
    behind).  After every step the ring and FRNN arrays it wrote must equal
    queries made afresh on the resulting mesh and positions."""
    import torch
    pts = ensemble_points(seed=11, sizes=batch_sizes())
    B = M.HostTriangulationBatch(pts)
    dt = M.DeviceTriangulation(B)
    dt.set_option(M.OPT_STAR_MIN_VERTICES, 0)
    dt.set_option(M.OPT_STAR_ENABLE, 1 if star else 0)
    rad = np.minimum(np.array([2.0 / np.sqrt(max(len(p), 1)) for p in pts], np.float32), np.float32(1.0 / 3.0))
    bg = M.GridCount2dBatch(B.vstart, rad)
    fresh = M.GridCount2dBatch(B.vstart, rad)
    nv, n = dt.m_pos_size, bg.numP
    fresh.build(dt.pos_tensor())
    cap = int(fresh.frnn_csr(torch.zeros(n + 1, dtype=torch.int32, device="cuda")) * 2) + 4096
    fo = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    fn = torch.zeros(cap, dtype=torch.int32, device="cuda")
    fo2 = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    fn2 = torch.zeros(cap, dtype=torch.int32, device="cuda")
    ro = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    rn = torch.zeros(dt.m_he_size, dtype=torch.int32, device="cuda")
    ro2 = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    rn2 = torch.zeros(dt.m_he_size, dtype=torch.int32, device="cuda")
    dups = 0
    for k in range(5):
        st = dt.timestep(M.MOTION_GAUSS, sigma, seed=5, step=k, ring=(ro, rn), bgrid=bg, frnn=(fo, fn))
        print("step", k, {x: st[x] for x in ("rounds", "fallbacks", "stalled", "duplicates")}, flush=True)
        dups += st["duplicates"] + st["stalled"]
        torch.cuda.synchronize()
        fresh.build(dt.pos_tensor())
        tot = fresh.frnn_csr(fo2, fn2)
        assert tot <= cap
        assert torch.equal(fo, fo2), "step %d: FRNN offsets differ" % k
        assert torch.equal(fn[:tot], fn2[:tot]), "step %d: FRNN lists differ" % k
        tr = dt.one_ring_csr(ro2, rn2)
        assert torch.equal(ro, ro2), "step %d: ring offsets differ" % k
        assert torch.equal(rn[:tr], rn2[:tr]), "step %d: rings differ" % k
        assert_at_targets(dt, st, "step %d" % k)
    _check_batch(oracle, B, dt, "after the timesteps")
    print("vertices left behind over the steps:", dups)
