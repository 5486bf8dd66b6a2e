"""Pins the oracle's triangulation algorithms (CPU only).

- Delaunay by brute force: every triangle's circumcircle is empty of every
  other vertex (exact, tie-broken incircle, itself pinned by the golden
  vectors), for small point sets in the box and on the torus.
- Counts fixed by Euler's formula: box T = 2n + 2, HE = 6n + 10 (the
  reference's corner layout); torus T = 2n, HE = 6n.
- The reference host path restated (Host_Triangulation.h:305-319 delonize,
  :417-471 untangle, :492-560 oneRing/getFRNN) on moved meshes: delonize
  reaches the unique triangulation, oneRing equals the adjacency read off the
  triangle list, getFRNN equals brute force.
"""
import numpy as np
import pytest

from tests.conftest import uniform_points


def _tris(he, t):
    he = np.asarray(he)
    out = []
    for h0 in np.asarray(t):
        h1 = he[h0, 0]
        h2 = he[h1, 0]
        out.append((he[h0, 1], he[h1, 1], he[h2, 1]))
    return out


def _empty_circles(O, pos, tris, periodic=False, near=0.3):
    """count (triangle, vertex) pairs with the vertex inside the circumcircle.
    On the torus only vertices within `near` (minimum image) of the triangle
    are tested: a minimum-image predicate is meaningful for a local
    neighbourhood, and these circumcircles are far smaller than `near`."""
    bad = 0
    pos = np.asarray(pos, np.float64)
    for a, b, c in tris:
        for p in range(len(pos)):
            if p in (a, b, c):
                continue
            if periodic:
                d = pos[p] - pos[a]
                d -= np.round(d)
                if np.hypot(*d) > near:
                    continue
            if O.incircle_sos(pos[a], pos[b], pos[c], pos[p], periodic) > 0:
                bad += 1
    return bad


@pytest.mark.parametrize("n,seed", [(12, 1), (40, 2), (120, 3)])
def test_box_dt_brute_force(oracle, n, seed):
    O = oracle
    xy = uniform_points(n, seed)
    T = O.delaunay(xy)
    pos, v, he, t = T.export()
    assert len(t) == 2 * n + 2 and len(he) == 6 * n + 10
    rc, msg = O.check_mesh(pos, v, he, t, periodic=False, require_ccw=True)
    assert rc == 0, msg
    assert _empty_circles(O, pos, _tris(he, t)) == 0


def test_torus_dt_brute_force(oracle):
    O = oracle
    xy = uniform_points(150, 9)
    tri = O.torus_delaunay(xy)
    assert len(tri) == 2 * len(xy)
    # the circumcircles are small (spacing ~0.08 << 1/2), so the minimum image suffices
    assert _empty_circles(O, xy, tri.tolist(), periodic=True) == 0
    e = set()
    for a, b, c in tri.tolist():
        e.update(((a, b), (b, c), (c, a)))
    assert all((q, p) in e for (p, q) in e) and len(e) == 6 * len(xy)


def test_dt_unique_over_insertion_orders(oracle):
    O = oracle
    xy = uniform_points(300, 5)
    base = O.canonical_triangles(*O.delaunay(xy).export()[2:])
    for s in range(3):
        perm = np.random.default_rng(s).permutation(len(xy))
        c = O.canonical_triangles(*O.delaunay(xy[perm]).export()[2:]).astype(np.int64)
        remap = np.concatenate([np.arange(4), 4 + perm])
        assert np.array_equal(base, O.canonical_from_list(remap[c]))


def _moved(O, n, seed, sigma, mode):
    xy = uniform_points(n, seed)
    T = O.HostTri(mode)
    T.add_delaunay_points(xy)
    pos, v, he, t = T.export()
    rng = np.random.default_rng(seed + 100)
    d = (sigma * rng.standard_normal((len(pos) - 4, 2))).astype(np.float32)
    newp = pos.copy()
    newp[4:] = np.clip(pos[4:] + d, 0.0, 1.0 - 2.0 ** -24)
    return T, newp


def test_delonize_restart_scan_and_worklist(oracle):
    """Host_Triangulation.h:305-319 (rescan from edge 0 after every flip) and
    its worklist restatement reach the same, unique, triangulation of the
    moved points when the move inverts nothing"""
    O = oracle
    for variant in ("restart", "worklist"):
        T, newp = _moved(O, 400, 11, 2e-4, O.HostTri.EXACT)
        T.set_positions(newp)
        dg = T.diagnostics()
        assert dg["non_ccw"] == 0 and dg["non_delaunay"] > 0
        flips = T.delonize() if variant == "restart" else T.delonize_worklist()
        assert flips > 0
        pos, v, he, t = T.export()
        assert O.count_non_delaunay(pos, he) == 0
        ref = O.canonical_triangles(*O.delaunay(pos[4:]).export()[2:])
        assert np.array_equal(O.canonical_triangles(he, t), ref), variant


def test_untangle_pass_reduces_inversions(oracle):
    """Host_Triangulation.h:417-471: one B/C pass flips creased edges; it never
    increases the number of inverted triangles on these inputs"""
    O = oracle
    T, newp = _moved(O, 2000, 12, 3e-3, O.HostTri.EXACT)
    T.set_positions(newp)
    before = T.diagnostics()["non_ccw"]
    assert before > 0
    flips = T.untangle()
    after = T.diagnostics()["non_ccw"]
    assert flips > 0 and after < before


def test_one_ring_equals_triangle_adjacency(oracle):
    O = oracle
    xy = uniform_points(500, 13)
    T = O.delaunay(xy)
    pos, v, he, t = T.export()
    adj = {i: set() for i in range(len(pos))}
    for a, b, c in _tris(he, t):
        for p, q in ((a, b), (b, c), (c, a)):
            adj[p].add(q)
            adj[q].add(p)
    for i in range(4, len(pos)):
        ring = T.one_ring(i)
        deg = len(adj[i])
        # Host_Triangulation.h:492-519 walks forward until it closes the
        # ring, then walks backward from the start as well: an interior
        # vertex gets its deg neighbours, then deg - 1 of them again in reverse
        assert set(ring) == adj[i]
        assert len(ring) == 2 * deg - 1
        fwd = ring[:deg]
        assert len(set(fwd)) == deg
        assert ring[deg:] == fwd[1:][::-1]
        # consecutive entries of the forward walk are adjacent (a rotation around i)
        for k in range(deg):
            assert fwd[(k + 1) % deg] in adj[fwd[k]]


@pytest.mark.parametrize("r", [0.02, 0.05, 0.1])
def test_graph_frnn_equals_brute_force(oracle, r):
    """getFRNN (Host_Triangulation.h:525-560) against the O(n^2) count, the
    check of tests/test_frnn.cu:61-72"""
    O = oracle
    xy = uniform_points(800, 14)
    T = O.delaunay(xy)
    pos = T.export()[0]
    d2 = ((pos[4:, None, :] - pos[None, 4:, :]) ** 2).sum(-1)
    for i in range(4, len(pos), 7):
        got = sorted(T.get_frnn(i, r))
        want = sorted(4 + j for j in np.nonzero(d2[i - 4] <= r * r)[0] if 4 + j != i)
        assert got == want, i
