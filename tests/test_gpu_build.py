"""The triangulation built on the GPU (mcl_tri_build) against the host build
and the oracle: same vertex numbering (the host build's Hilbert keys), same
positions, the same canonical triangle set, a valid mesh with every edge
exactly locally Delaunay -- on uniform, clustered, tied-coordinate and
near-degenerate (lattice + 1e-7 jitter) point sets, and the tiny cases.
Point sets that lie on one line are reported (MCL_ERR_UNRESOLVED)."""
import numpy as np
import pytest

import mcleap_amd as M
from mcleap_amd.workloads import unique_points

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(5)
    out = {
        "n1": rng.random((1, 2)),
        "n2": rng.random((2, 2)),
        "n3": rng.random((3, 2)),
        "n17": rng.random((17, 2)),
        "uniform_20k": rng.random((20000, 2)),
        "clustered_5k": 0.5 + 0.02 * rng.standard_normal((5000, 2)),
        "ties_grid": np.floor(rng.random((3000, 2)) * 40) / 40,
    }
    g = np.stack(np.meshgrid(np.arange(60), np.arange(60), indexing="ij"), -1).reshape(-1, 2) / 60.0
    out["lattice_jitter"] = np.clip(g + rng.uniform(-1e-7, 1e-7, g.shape), 0.0, 1.0 - 2.0 ** -24)
    return {k: unique_points(np.asarray(v, np.float32)) for k, v in out.items()}


@pytest.mark.parametrize("name", list(_cases()))
def test_device_build_equals_host_build(oracle, name):
    O = oracle
    xy = _cases()[name]
    print(name, len(xy), flush=True)
    H = M.HostTriangulation()
    q = xy.copy()
    H.addDelaunayPoints(q)
    D = M.DeviceTriangulation.from_points(xy)
    h = D.h_triangulation
    assert np.array_equal(h.perm, H.perm)
    assert np.array_equal(h.m_pos, H.m_pos)
    rc, msg = O.check_mesh(h.m_pos, h.m_v, h.m_he, h.m_t, periodic=False, require_ccw=True)
    assert rc == 0, msg
    assert O.count_non_delaunay(h.m_pos, h.m_he) == 0
    assert np.array_equal(O.canonical_triangles(h.m_he, h.m_t), O.canonical_triangles(H.m_he, H.m_t))
    # and the device object steps on as any other
    st = D.step(M.MOTION_GAUSS, 1e-3, seed=2, step=0)
    assert D.count_inverted() == 0 and D.count_non_delaunay() == 0, st


def test_device_build_collinear_input_is_reported():
    """all points on one line: moving the row positions onto it stalls (a
    vertex cannot leave a collinear link), and the build says so rather than
    returning a mesh off its points; mcl_hmesh_build handles such input"""
    xy = np.stack([np.full(300, 0.25), np.random.default_rng(2).random(300)], 1).astype(np.float32)
    with pytest.raises(M.MclError) as e:
        M.DeviceTriangulation.from_points(xy)
    assert e.value.code == 6  # MCL_ERR_UNRESOLVED


def test_device_build_rejects_duplicates():
    xy = np.random.default_rng(1).random((100, 2)).astype(np.float32)
    xy[57] = xy[12]
    with pytest.raises(M.MclError):
        M.DeviceTriangulation.from_points(xy)


@pytest.mark.parametrize("n", [200, 20000])
def test_device_build_torus_equals_host_build(oracle, n):
    """the torus: rows closed into loops, strips into annuli; the same
    vertex numbering, positions and canonical triangles as the host's torus
    build, a valid periodic mesh with every edge exactly locally Delaunay"""
    O = oracle
    xy = unique_points(np.random.default_rng(n).random((n, 2)).astype(np.float32))
    xy[xy >= 1.0] = 0.0
    H = M.HostTriangulation(periodic=True)
    q = xy.copy()
    H.addDelaunayPoints(q)
    D = M.DeviceTriangulation.from_points(xy, periodic=True)
    h = D.h_triangulation
    assert h.periodic and np.array_equal(h.perm, H.perm) and np.array_equal(h.m_pos, H.m_pos)
    rc, msg = O.check_mesh(h.m_pos, h.m_v, h.m_he, h.m_t, periodic=True, require_ccw=True)
    assert rc == 0, msg
    assert O.count_non_delaunay(h.m_pos, h.m_he, periodic=True) == 0
    assert np.array_equal(O.canonical_triangles(h.m_he, h.m_t), O.canonical_triangles(H.m_he, H.m_t))
    st = D.step(M.MOTION_SHEAR, 2e-4, amp=1e-3, seed=3, step=0)
    assert D.count_inverted() == 0 and D.count_non_delaunay() == 0, st


def test_device_build_torus_too_few_points():
    xy = np.random.default_rng(1).random((10, 2)).astype(np.float32)
    with pytest.raises(M.MclError):
        M.DeviceTriangulation.from_points(xy, periodic=True)
