"""CPU tests of the product's host side and of the C-ABI library surface.

- libmcleap_hip.so loads without a GPU and exports every function that
  include/mcleap_hip.h declares (no compute calls are made here).
- The host builder behind HostTriangulation.addDelaunayPoints (C++ in the
  library, Host_Triangulation.h:296-300 semantics, exact predicates) gives
  the same canonical triangle set as the oracle's restated incremental
  Lawson build, in the box with the four +-1e7 corners and on the torus.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import mcleap_amd as M
from tests.conftest import uniform_points

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(mcl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(M._lib.LIB_PATH)
    names = _declared("mcleap_hip.h") + _declared("mcleap_hip_diag.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table binds exactly what the header declares
    assert set(M._lib.SIGNATURES) <= set(names)


def test_version_string():
    v = M.lib().mcl_version()
    assert v and b"gfx950" in v


def test_corner_layout_matches_reference():
    """Host_Triangulation.h:107-137: 4 corners, 2 triangles, 10 half-edges"""
    H = M.HostTriangulation()
    assert H.m_pos.shape == (4, 2) and H.m_he.shape == (10, 4) and H.m_t.shape == (2,)
    assert set(map(tuple, H.m_pos.tolist())) == {(-1e7, -1e7), (1e7, -1e7), (1e7, 1e7), (-1e7, 1e7)}
    # outer face: 4 half-edges with t = op = -1
    outer = H.m_he[H.m_he[:, 2] < 0]
    assert len(outer) == 4 and np.all(outer[:, 3] == -1)


@pytest.mark.parametrize("n,seed", [(50, 1), (2000, 2), (20000, 3)])
def test_host_builder_box_equals_oracle(oracle, n, seed):
    O = oracle
    xy = uniform_points(n, seed)
    H = M.HostTriangulation()
    H.addDelaunayPoints(xy)
    assert H.m_pos.shape[0] == len(xy) + 4
    # the caller's array was reordered in place to the vertex order
    assert np.array_equal(H.m_pos[4:], xy)
    rc, msg = O.check_mesh(H.m_pos, H.m_v, H.m_he, H.m_t, periodic=False, require_ccw=True)
    assert rc == 0, msg
    assert H.m_he.shape[0] == 10 + 6 * len(xy) and H.m_t.shape[0] == 2 + 2 * len(xy)
    assert O.count_non_delaunay(H.m_pos, H.m_he) == 0
    ref = O.delaunay(H.m_pos[4:])
    _, _, h2, t2 = ref.export()
    assert np.array_equal(O.canonical_triangles(H.m_he, H.m_t), O.canonical_triangles(h2, t2))


@pytest.mark.parametrize("n,seed", [(300, 4), (5000, 5)])
def test_host_builder_torus_equals_oracle(oracle, n, seed):
    O = oracle
    xy = uniform_points(n, seed)
    H = M.HostTriangulation(periodic=True)
    H.addDelaunayPoints(xy)
    assert H.m_t.shape[0] == 2 * len(xy) and H.m_he.shape[0] == 6 * len(xy)
    rc, msg = O.check_mesh(H.m_pos, H.m_v, H.m_he, H.m_t, periodic=True, require_ccw=True)
    assert rc == 0, msg
    assert O.count_non_delaunay(H.m_pos, H.m_he, periodic=True) == 0
    tri = O.torus_delaunay(H.m_pos)
    assert np.array_equal(O.canonical_triangles(H.m_he, H.m_t), O.canonical_from_list(tri))


def test_host_builder_rejects_duplicates():
    xy = uniform_points(100, 6)
    xy = np.concatenate([xy, xy[:1]])
    H = M.HostTriangulation()
    with pytest.raises(M.MclError):
        H.addDelaunayPoints(xy)


def test_cocircular_lattice_is_unique_under_the_tie_rule(oracle):
    """an exact integer lattice is fully degenerate (every square cocircular);
    the symbolic perturbation still defines one triangulation, and both
    builders (different insertion orders) must produce it"""
    O = oracle
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), indexing="ij"), -1).reshape(-1, 2) / 16.0
    xy = g.astype(np.float32)
    H = M.HostTriangulation()
    H.addDelaunayPoints(xy)
    ref = O.delaunay(H.m_pos[4:])
    _, _, h2, t2 = ref.export()
    assert np.array_equal(O.canonical_triangles(H.m_he, H.m_t), O.canonical_triangles(h2, t2))
    # a shuffled insertion order gives the same set
    perm = np.random.default_rng(0).permutation(len(xy))
    ref2 = O.delaunay(H.m_pos[4:][perm])
    _, _, h3, t3 = ref2.export()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    c3 = O.canonical_triangles(h3, t3).astype(np.int64)
    remap = np.concatenate([np.arange(4), 4 + perm])  # oracle vertex 4+k holds point perm[k]
    c3 = O.canonical_from_list(remap[c3])
    assert np.array_equal(O.canonical_triangles(H.m_he, H.m_t), c3)


def _dt_equal(O, H, tag):
    """H (box) is a valid mesh equal to the oracle's exact DT of its positions"""
    rc, msg = O.check_mesh(H.m_pos, H.m_v, H.m_he, H.m_t, periodic=False, require_ccw=True)
    assert rc == 0, "%s: %s" % (tag, msg)
    assert O.count_non_delaunay(H.m_pos, H.m_he) == 0, tag
    ref = O.delaunay(H.m_pos[4:])
    _, _, h2, t2 = ref.export()
    assert np.array_equal(O.canonical_triangles(H.m_he, H.m_t), O.canonical_triangles(h2, t2)), tag


def test_move_flipflop_host(oracle):
    """moveFlipflop (Host_Triangulation.h:77, a stub in the reference): a
    vertex removed and re-inserted two spacings away keeps every triangle
    strictly CCW and lands exactly at pos + d; moveFlipflopDelaunay then
    equals the oracle's exact Delaunay triangulation of the positions"""
    O = oracle
    n = 2000
    H = M.HostTriangulation()
    H.addDelaunayPoints(uniform_points(n, 8))
    rng = np.random.default_rng(4)
    sp = 1.0 / np.sqrt(n)
    moved = 0
    for k in range(60):
        v = int(rng.integers(4, n + 4))
        d = (2.0 * sp * rng.standard_normal(2)).astype(np.float32)
        want = H.m_pos[v] + d
        if np.any(want <= 0) or np.any(want >= 1):
            continue
        delaunay = k >= 30  # first moveFlipflop alone, then (from a Delaunay mesh) with Lawson flips
        if k == 30:
            assert H.delonize()
            _dt_equal(O, H, "moveFlipflop x30 + delonize")
        ok = (H.moveFlipflopDelaunay if delaunay else H.moveFlipflop)(v, d)
        assert ok, k
        moved += 1
        assert np.array_equal(H.m_pos[v], want.astype(np.float32)), k
        rc, msg = O.check_mesh(H.m_pos, H.m_v, H.m_he, H.m_t, periodic=False, require_ccw=True)
        assert rc == 0, "step %d: %s" % (k, msg)
        if delaunay:
            _dt_equal(O, H, "moveFlipflopDelaunay %d" % k)
    assert moved > 40
    assert H.moveUntanglingDelaunay(5, np.float32([1e-4, -1e-4]))
    _dt_equal(O, H, "moveUntanglingDelaunay")


def test_move_flipflop_degenerate_link(oracle):
    """a vertex at the centre of a square of four neighbours (v on the
    diagonal of every ear of its link) is removed by the proper-ear rule and
    re-inserted outside its old star; also a target on another vertex
    (refused) and on an edge (2 -> 4 insertion)"""
    O = oracle
    c = np.float32(0.5)
    ring = [(0.5 + 0.3 * np.cos(a), 0.5 + 0.3 * np.sin(a)) for a in np.linspace(0, 2 * np.pi, 24, endpoint=False)]
    pts = np.array([(c, c), (0.4, 0.5), (0.6, 0.5), (0.5, 0.4), (0.5, 0.6)] + ring, np.float32)
    H = M.HostTriangulation()
    H.addDelaunayPoints(pts.copy())
    vid = {tuple(p): i for i, p in enumerate(H.m_pos.tolist())}
    v = vid[(0.5, 0.5)]
    deg = np.sum(H.m_he[:, 1] == v)
    assert deg == 4
    assert H.moveFlipflop(v, np.float32([0.17, 0.06]))
    assert np.array_equal(H.m_pos[v], np.float32([0.5, 0.5]) + np.float32([0.17, 0.06]))
    rc, msg = O.check_mesh(H.m_pos, H.m_v, H.m_he, H.m_t, periodic=False, require_ccw=True)
    assert rc == 0, msg
    assert H.delonize()
    _dt_equal(O, H, "after the degenerate removal")
    # onto another vertex: refused, mesh unchanged
    before = [a.copy() for a in (H.m_pos, H.m_v, H.m_he, H.m_t)]
    u = vid[(np.float32(0.4), np.float32(0.5))]
    assert not H.moveFlipflop(v, H.m_pos[u] - H.m_pos[v])
    for a, b in zip(before, (H.m_pos, H.m_v, H.m_he, H.m_t)):
        assert np.array_equal(a, b)
    # onto the midpoint of the edge between two lattice neighbours (exact in FP32)
    w = vid[(np.float32(0.5), np.float32(0.4))]
    mid = np.float32([0.45, 0.45])
    assert H.moveFlipflopDelaunay(w, mid - H.m_pos[w])
    assert np.array_equal(H.m_pos[w], mid)
    _dt_equal(O, H, "after the on-edge insertion")


def test_tuning_table_loaded_from_env(tmp_path):
    """the autotuner's table is read at load (MCL_TUNING): a well-formed file
    is reported by mcl_tuning_path, a missing one or MCL_TUNING=off is not,
    and the Python lookup picks the nearest entry of the same domain"""
    import json
    import subprocess
    import sys
    from mcleap_amd import tuning as T
    table = {"entries": [{"domain": "torus", "log2_nv": 20, "graph": 1, "star_head_passes": 3, "grid_block": 128,
                          "frnn_cells_per_radius": 2, "tuned_set_kept": True, "note": None},
                         {"domain": "box", "log2_nv": 14, "tail_threshold": 4096, "tuned_set_kept": False}]}
    path = tmp_path / "t.json"
    path.write_text(json.dumps(table))
    code = "import mcleap_amd as M; print(M.lib().mcl_tuning_path().decode())"
    shipped = os.path.join(ROOT, "mcleap_amd", "tuning.json")
    for env, want in ((str(path), str(path)), ("off", ""), (str(tmp_path / "none.json"), ""), (None, shipped)):
        e = dict(os.environ)
        e.pop("MCL_TUNING", None)
        if env is not None:
            e["MCL_TUNING"] = env
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, env=e,
                             timeout=120)
        assert out.returncode == 0, out.stderr
        # every entry parsed (record fields such as booleans are skipped, not errors)
        assert "malformed" not in out.stderr, (env, out.stderr)
        assert os.path.realpath(out.stdout.strip() or "/nonexistent") == os.path.realpath(want) if want else \
            out.stdout.strip() == "", (env, out.stdout)
    assert T.lookup("torus", 1 << 20, str(path))["star_head_passes"] == 3
    assert T.lookup("torus", 1 << 22, str(path)) is not None and T.lookup("torus", 1 << 23, str(path)) is None
    assert T.cells_per_radius("torus", 1 << 20, str(path)) == 2
    assert T.cells_per_radius("box", 16000, str(path)) == 1
    assert T.lookup("ensemble", 1 << 20, str(path)) is None
