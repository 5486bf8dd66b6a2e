"""GPU one-ring and closest-neighbour queries against the oracle.

- oneRing<K> / oneRing2<K> layouts (kernels.cuh:54-121): count in slot 0,
  at most K-1 neighbours, CW rotation from v.he; the neighbour set equals the
  adjacency read off the downloaded triangle list, and consecutive entries
  turn clockwise around the vertex (exact orientation).
- CSR one-ring: no cap, same order, total = number of half-edges.
- closestNeighbors (kernels.cuh:129-146; tests/test_closest_neighbors.cu):
  mt19937(3), uniform_real_distribution<float>(-1000, 1000), 10 / 100 / 1000
  points; the closest ring neighbour equals the brute-force closest point
  (same FP32 squared distance, first minimum).  The four +-1e7 corners are
  left out: at that distance every FP32 squared distance to a unit-scale
  point rounds to a handful of values, so the first minimum depends on the
  scan order, in the reference as here.
"""
import numpy as np
import pytest

import mcleap_amd as M
from tests.conftest import uniform_points

pytestmark = pytest.mark.gpu


def _adjacency(he, t, nv):
    adj = [set() for _ in range(nv)]
    for h0 in t:
        h1 = he[h0, 0]
        h2 = he[h1, 0]
        a, b, c = he[h0, 1], he[h1, 1], he[h2, 1]
        for p, q in ((a, b), (b, c), (c, a)):
            adj[p].add(q)
            adj[q].add(p)
    return adj


def _mesh(n, seed, periodic=False):
    xy = uniform_points(n, seed)
    H = M.HostTriangulation(periodic=periodic)
    H.addDelaunayPoints(xy)
    return M.DeviceTriangulation(H)


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("K", [10, 5])
def test_legacy_ring_layouts(oracle, layout, K):
    import torch
    dt = _mesh(3000, 21)
    pos, v, he, t = dt.download()
    nv = len(pos)
    ring = torch.full((nv * K,), -7, dtype=torch.int32, device="cuda")
    ok = (dt.oneRing if layout == 0 else dt.oneRing2)(K, ring)
    r = ring.cpu().numpy()
    adj = _adjacency(he, t, nv)
    # points with more than K - 1 neighbours are counted (the reference
    # truncates them silently) and the call reports it; the corners' fans
    # are truncated without being counted
    over = sum(len(a) > K - 1 for a in adj[4:])
    assert dt.ring_overflow == over and ok == (over == 0)
    if K == 5:
        assert over > 0
    for i in range(4, nv):
        cnt = r[i] if layout == 0 else r[K * i]
        ids = [r[(c + 1) * nv + i] if layout == 0 else r[(c + 1) + K * i] for c in range(cnt)]
        assert cnt == min(len(adj[i]), K - 1)
        assert set(ids) <= adj[i] and len(set(ids)) == cnt
        if cnt == len(adj[i]):
            assert set(ids) == adj[i]
        for k in range(cnt - 1):  # clockwise turn around i
            assert oracle.orient_exact(pos[i], pos[ids[k]], pos[ids[k + 1]]) < 0


@pytest.mark.parametrize("periodic", [False, True])
def test_csr_ring(oracle, periodic):
    import torch
    dt = _mesh(5000, 22, periodic)
    pos, v, he, t = dt.download()
    nv = len(pos)
    off = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    tot = dt.one_ring_csr(off)
    assert tot == len(he)
    nbr = torch.zeros(tot, dtype=torch.int32, device="cuda")
    dt.one_ring_csr(off, nbr)
    off, nbr = off.cpu().numpy(), nbr.cpu().numpy()
    adj = _adjacency(he, t, nv)
    for i in range(nv):
        ids = nbr[off[i]:off[i + 1]].tolist()
        assert set(ids) == adj[i] and len(ids) == len(adj[i])
    # the asynchronous form writes the same arrays
    off2 = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    nbr2 = torch.zeros(len(he), dtype=torch.int32, device="cuda")
    dt.one_ring_csr_async(off2, nbr2)
    torch.cuda.synchronize()
    assert np.array_equal(off2.cpu().numpy(), off) and np.array_equal(nbr2.cpu().numpy(), nbr)


def test_csr_ring_after_star_steps(oracle):
    """After torus steps (which keep the triangulation as per-vertex stars and
    rebuild the half-edge arrays only when they are read), the CSR one-ring
    equals, entry for entry and in order, the clockwise walk from v.he over
    the downloaded half-edge mesh."""
    import torch
    dt = _mesh(20000, 23, True)
    dt.set_option(M.OPT_STAR_MIN_VERTICES, 0)
    dt.set_option(M.OPT_STAR_ENABLE, 1)
    for k in range(2):
        st = dt.step(M.MOTION_SHEAR, 2e-4, amp=1e-3, seed=5, step=k)
        assert st["fallbacks"] == 0
    nv = dt.m_pos_size
    off = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    nbr = torch.zeros(dt.m_he_size, dtype=torch.int32, device="cuda")
    dt.one_ring_csr_async(off, nbr)
    torch.cuda.synchronize()
    off, nbr = off.cpu().numpy(), nbr.cpu().numpy()
    pos, v, he, t = dt.download()
    for i in range(nv):
        h0 = h = int(v[i, 1])
        walk = []
        while True:
            walk.append(int(he[h ^ 1, 1]))
            h = int(he[h ^ 1, 0])
            if h == h0:
                break
        assert nbr[off[i]:off[i + 1]].tolist() == walk, i


@pytest.mark.parametrize("n", [10, 100, 1000])
def test_closest_neighbors_reference_case(oracle, n):
    import torch
    d = oracle.mt_uniform_floats(3, -1000.0, 1000.0, 2 * n)
    xy = d.reshape(-1, 2)
    H = M.HostTriangulation()
    H.addDelaunayPoints(xy)
    dt = M.DeviceTriangulation(H)
    pos = dt.download_positions()
    nv = len(pos)
    off = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
    tot = dt.one_ring_csr(off)
    nbr = torch.zeros(tot, dtype=torch.int32, device="cuda")
    dt.one_ring_csr(off, nbr)
    out = torch.zeros(nv, dtype=torch.int32, device="cuda")
    dt.closestNeighbors(off, nbr, out)
    got = out.cpu().numpy()
    for i in range(4, nv):
        dx = (pos[i, 0] - pos[:, 0]).astype(np.float32)
        dy = (pos[i, 1] - pos[:, 1]).astype(np.float32)
        d2 = dx * dx + dy * dy
        d2[i] = np.inf
        best = np.flatnonzero(d2 == d2.min())
        assert got[i] in best, (i, got[i], best)


@pytest.mark.parametrize("layout", [0, 1])
def test_closest_over_legacy_ring(oracle, layout):
    """closestNeighbors<K> over oneRing (SoA) and closestNeighbors2<K> over oneRing2 (AoS)"""
    import torch
    d = oracle.mt_uniform_floats(3, -1000.0, 1000.0, 2000)
    H = M.HostTriangulation()
    H.addDelaunayPoints(d.reshape(-1, 2))
    dt = M.DeviceTriangulation(H)
    pos = dt.download_positions()
    nv, K = len(pos), 100
    ring = torch.zeros(nv * K, dtype=torch.int32, device="cuda")
    (dt.oneRing if layout == 0 else dt.oneRing2)(K, ring)
    out = torch.zeros(nv, dtype=torch.int32, device="cuda")
    dt.closestNeighbors_ring(K, ring, out, layout)
    got = out.cpu().numpy()
    for i in range(4, nv):
        dx = (pos[i, 0] - pos[:, 0]).astype(np.float32)
        dy = (pos[i, 1] - pos[:, 1]).astype(np.float32)
        d2 = dx * dx + dy * dy
        d2[i] = np.inf
        assert got[i] in np.flatnonzero(d2 == d2.min())


@pytest.mark.parametrize("radius", [5.0, 10.0, 20.0, 30.0, 40.0, 50.0])
def test_graph_frnn_reference_case(oracle, radius):
    """tests/test_frnn.cu: mt19937(3), 1000 points in [-1000, 1000)^2, K = 20,
    M = 100; the count equals the brute-force count within r (the
    reference's check) and the id set equals the oracle's getFRNN walk
    (Host_Triangulation.h:525-560) on the same mesh"""
    import torch
    d = oracle.mt_uniform_floats(3, -1000.0, 1000.0, 2000)
    H = M.HostTriangulation()
    H.addDelaunayPoints(d.reshape(-1, 2))
    dt = M.DeviceTriangulation(H)
    pos, v, he, t = dt.download()
    nv, K, Mx = len(pos), 20, 100
    T = oracle.HostTri(oracle.HostTri.EXACT)
    T.load(pos, v, he, t)
    for layout in (0, 1):
        ring = torch.zeros(nv * K, dtype=torch.int32, device="cuda")
        (dt.oneRing if layout == 0 else dt.oneRing2)(K, ring)
        out = torch.full((nv * Mx,), -1, dtype=torch.int32, device="cuda")
        dt.getFRNN(K, Mx, radius, ring, out, layout)
        o = out.cpu().numpy()
        rr = np.float32(radius) * np.float32(radius)
        for i in range(4, nv):
            cnt = o[i] if layout == 0 else o[Mx * i]
            ids = [o[(k + 1) * nv + i] if layout == 0 else o[(k + 1) + Mx * i] for k in range(min(cnt, Mx - 1))]
            dx = (pos[i, 0] - pos[4:, 0]).astype(np.float32)
            dy = (pos[i, 1] - pos[4:, 1]).astype(np.float32)
            brute = int(((dx * dx + dy * dy) <= rr).sum()) - 1
            assert cnt == brute, (layout, i, cnt, brute)
            assert sorted(ids) == sorted(T.get_frnn(i, radius))
