"""Host build of the batched ensemble (mcl_hmesh_build_batch): the union
mesh holds every instance's own exact triangulation with its ids offset, the
corners of all instances first."""
import numpy as np

import mcleap_amd as M
from mcleap_amd.workloads import ensemble_points, ensemble_sizes


def batch_sizes():
    """an empty instance, tiny ones, and log-uniform sizes (24 instances)"""
    return np.concatenate([[0, 1, 2, 3, 17], ensemble_sizes(19, seed=11, nmin=64, nmax=3000)])


def test_batch_build_matches_single(oracle):
    pts = ensemble_points(seed=11, sizes=batch_sizes())
    B = M.HostTriangulationBatch(pts, nthreads=3)
    assert B.ninst == len(pts) and B.vstart[0] == 4 * B.ninst
    assert B.vstart[-1] == len(B.m_pos) and B.hestart[-1] == len(B.m_he) and B.tstart[-1] == len(B.m_t)
    off = 0
    for i, p in enumerate(pts):
        H = M.HostTriangulation()
        q = p.copy()
        H.addDelaunayPoints(q)
        vs = B.instance_vertices(i)
        hs, ts = B.hestart[i], B.tstart[i]
        assert np.array_equal(B.m_pos[vs], H.m_pos)
        g = B.m_he[hs:B.hestart[i + 1]]
        loc = H.m_he
        assert np.array_equal(g[:, 0] - hs, loc[:, 0])
        assert np.array_equal(g[:, 1], vs[loc[:, 1]])
        assert np.array_equal(g[:, 2], np.where(loc[:, 2] >= 0, loc[:, 2] + ts, loc[:, 2]))
        assert np.array_equal(g[:, 3], np.where(loc[:, 3] >= 0, vs[np.maximum(loc[:, 3], 0)], loc[:, 3]))
        assert np.array_equal(B.m_t[ts:B.tstart[i + 1]] - hs, H.m_t)
        assert np.array_equal(B.m_v[vs, 1] - hs, H.m_v[:, 1]) and np.array_equal(B.m_v[:, 0], np.arange(len(B.m_v)))
        seg = B.perm[B.vstart[i] - 4 * B.ninst:B.vstart[i + 1] - 4 * B.ninst]
        assert np.array_equal(seg - off, H.perm)
        # and the instance is the oracle's exact DT of its points
        ref = oracle.delaunay(H.m_pos[4:])
        p2, v2, h2, t2 = ref.export()
        assert np.array_equal(oracle.canonical_triangles(H.m_he, H.m_t), oracle.canonical_triangles(h2, t2))
        off += len(p)


def test_batch_build_rejects_bad_input():
    bad = [np.array([[0.1, 0.1], [0.1, 0.1]], np.float32)]  # a duplicate point
    try:
        M.HostTriangulationBatch([np.zeros((3, 2), np.float32) + [[0.2, 0.3]], *bad])
    except M.MclError as e:
        assert "3" in str(e) or "DUPLICATE" in str(e).upper()
    else:
        raise AssertionError("duplicate points accepted")


def test_state_file_round_trip(tmp_path, oracle):
    """mcl_state_save / info / load: the arrays come back bit for bit; a
    truncated or corrupted file, or another file, is rejected"""
    xy = np.random.default_rng(3).random((500, 2)).astype(np.float32)
    H = M.HostTriangulation()
    H.addDelaunayPoints(xy)
    path = tmp_path / "tri.mcl"
    H.save_state(path)
    G = M.HostTriangulation.load_state(path)
    assert not G.periodic
    for a, b in ((H.m_pos, G.m_pos), (H.m_v, G.m_v), (H.m_he, G.m_he), (H.m_t, G.m_t)):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    raw = path.read_bytes()
    assert raw[:8] == b"MCLTRI01" and len(raw) == 40 + 8 * len(H.m_pos) + 8 * len(H.m_v) + 16 * len(H.m_he) + \
        4 * len(H.m_t) + 8
    bad = bytearray(raw)
    bad[40 + 123] ^= 0x10
    (tmp_path / "bad.mcl").write_bytes(bytes(bad))
    (tmp_path / "short.mcl").write_bytes(raw[:-9])
    (tmp_path / "long.mcl").write_bytes(raw + b"\0")
    (tmp_path / "other.mcl").write_bytes(b"not a state file at all, no magic here")
    for name in ("bad.mcl", "short.mcl", "long.mcl", "other.mcl", "missing.mcl"):
        try:
            M.HostTriangulation.load_state(tmp_path / name)
        except M.MclError:
            pass
        else:
            raise AssertionError(name + " accepted")
    # a torus state keeps its flag
    T = M.HostTriangulation(periodic=True)
    T.addDelaunayPoints(xy)
    T.save_state(tmp_path / "torus.mcl")
    assert M.HostTriangulation.load_state(tmp_path / "torus.mcl").periodic


def test_state_file_with_bad_indices_rejected_before_the_device(tmp_path):
    """a state file whose checksum is valid but whose arrays break the mesh
    structure (indices out of range, next^3 != id, v.he / t.he pointing
    elsewhere) is refused by mcl_tri_load and mcl_tri_create with
    MCL_ERR_INPUT on the host, before anything reaches the device"""
    import ctypes
    xy = np.random.default_rng(5).random((300, 2)).astype(np.float32)
    H = M.HostTriangulation()
    H.addDelaunayPoints(xy)
    nhe, nv, nt = len(H.m_he), len(H.m_v), len(H.m_t)
    inner = int(np.flatnonzero(H.m_he[:, 2] >= 0)[7])

    def variant(fn):
        pos, v, he, t = H.m_pos.copy(), H.m_v.copy(), H.m_he.copy(), H.m_t.copy()
        fn(pos, v, he, t)
        return pos, v, he, t

    cases = {
        "next out of range": lambda p, v, he, t: he.__setitem__((3, 0), nhe + 5),
        "origin out of range": lambda p, v, he, t: he.__setitem__((3, 1), -2),
        "triangle out of range": lambda p, v, he, t: he.__setitem__((inner, 2), nt),
        "opposite out of range": lambda p, v, he, t: he.__setitem__((inner, 3), nv + 1),
        "next^3": lambda p, v, he, t: he.__setitem__((inner, 0), inner),
        "v.he": lambda p, v, he, t: v.__setitem__((9, 1), int(v[10, 1])),
        "t.he": lambda p, v, he, t: t.__setitem__(4, int(t[5])),
        "nan position": lambda p, v, he, t: p.__setitem__((6, 0), np.nan),
    }
    L = M.lib()
    for name, fn in cases.items():
        pos, v, he, t = variant(fn)
        path = tmp_path / "bad.mcl"
        rc = L.mcl_state_save(str(path).encode(), 0, pos.ctypes.data_as(ctypes.c_void_p), len(pos),
                              v.ctypes.data_as(ctypes.c_void_p), nv,
                              he.ctypes.data_as(ctypes.c_void_p), nhe, t.ctypes.data_as(ctypes.c_void_p), nt)
        assert rc == 0, name
        h = ctypes.c_void_p()
        assert L.mcl_tri_load(str(path).encode(), ctypes.byref(h)) == 2, name
        rc = L.mcl_tri_create(pos.ctypes.data_as(ctypes.c_void_p), len(pos), v.ctypes.data_as(ctypes.c_void_p), nv,
                              he.ctypes.data_as(ctypes.c_void_p), nhe, t.ctypes.data_as(ctypes.c_void_p), nt, 0,
                              ctypes.byref(h))
        assert rc == 2, name
    # the unmodified arrays pass the check (here: no device, so the
    # allocation that follows reports that instead)
    h = ctypes.c_void_p()
    rc = L.mcl_tri_create(H.m_pos.ctypes.data_as(ctypes.c_void_p), nv, H.m_v.ctypes.data_as(ctypes.c_void_p), nv,
                          H.m_he.ctypes.data_as(ctypes.c_void_p), nhe, H.m_t.ctypes.data_as(ctypes.c_void_p), nt, 0,
                          ctypes.byref(h))
    assert rc != 2
    if rc == 0:
        L.mcl_tri_destroy(h)
