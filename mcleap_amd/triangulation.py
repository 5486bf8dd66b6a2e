"""Python mirror of the reference's host-side interface, over the C ABI.

Class and method names follow the reference's C++ types so tests and scripts
read like the reference's own:
  HostTriangulation    Host_Triangulation.h:24-98
  DeviceTriangulation  Device_Triangulation.h:128-223
  GridCount2d          gridCount2d.h:23-48
Host arrays are numpy AoS records in the reference layouts
(m_pos float32 (V,2), m_v int32 (V,2) = {pos_index, he},
m_he int32 (HE,4) = {next, v, t, op}, m_t int32 (T,) = {he}).
Device buffers are torch tensors on the GPU; every device call goes to
libmcleap_hip.so and raises on failure (there is no CPU fallback).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

_vp = ctypes.c_void_p


def _np_ptr(a):
    return a.ctypes.data_as(_vp)


def _dev_ptr(x):
    """device pointer of a torch tensor (or an int address)"""
    if x is None:
        return None
    if isinstance(x, int):
        return _vp(x)
    if not x.is_cuda:
        raise ValueError("expected a device tensor")
    if not x.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return _vp(x.data_ptr())


class HostTriangulation:
    """Host triangulation (Host_Triangulation.h:24-98).

    The constructor holds the 4 corner vertices at +-1e7 and 2 triangles
    (Host_Triangulation.h:107-137).  addDelaunayPoints builds the exact
    Delaunay triangulation in libmcleap_hip (Hilbert insertion order) and,
    like the reference, reorders the caller's point array in place so that
    vertex 4 + k holds points[k].  periodic=True builds the unit-torus
    triangulation (no corners; vertex k holds points[k]).
    """

    def __init__(self, periodic=False):
        self.periodic = bool(periodic)
        self._points = np.zeros((0, 2), np.float32)
        self.perm = np.zeros(0, np.int32)
        if self.periodic:
            self.m_pos = np.zeros((0, 2), np.float32)
            self.m_v = np.zeros((0, 2), np.int32)
            self.m_he = np.zeros((0, 4), np.int32)
            self.m_t = np.zeros(0, np.int32)
        else:
            self._build(np.zeros((0, 2), np.float32))

    def _build(self, pts):
        L = lib()
        h = _vp()
        pts = np.ascontiguousarray(pts, np.float32)
        check(L.mcl_hmesh_build(_np_ptr(pts), len(pts), int(self.periodic), ctypes.byref(h)), "mcl_hmesh_build")
        try:
            sz = [ctypes.c_int() for _ in range(4)]
            check(L.mcl_hmesh_sizes(h, *[ctypes.byref(s) for s in sz]), "mcl_hmesh_sizes")
            npos, nv, nhe, nt = (s.value for s in sz)
            self.m_pos = np.zeros((npos, 2), np.float32)
            self.m_v = np.zeros((nv, 2), np.int32)
            self.m_he = np.zeros((nhe, 4), np.int32)
            self.m_t = np.zeros(nt, np.int32)
            perm = np.zeros(len(pts), np.int32)
            check(L.mcl_hmesh_export(h, _np_ptr(self.m_pos), _np_ptr(self.m_v), _np_ptr(self.m_he),
                                     _np_ptr(self.m_t), _np_ptr(perm)), "mcl_hmesh_export")
        finally:
            L.mcl_hmesh_free(h)
        self.perm = perm
        self._points = pts[perm]

    def addDelaunayPoints(self, points):
        """Host_Triangulation.h:296-300.  Reorders `points` in place (numpy)."""
        pts = np.asarray(points, np.float32).reshape(-1, 2)
        allpts = np.concatenate([self._points, pts]) if len(self._points) else pts
        self._build(allpts)
        if isinstance(points, np.ndarray) and points.shape == pts.shape and len(self._points) == len(pts):
            points[...] = self._points
        return True

    def num_points(self):
        return len(self._points)

    def _arrays(self):
        for a in ("m_pos", "m_v", "m_he", "m_t"):
            setattr(self, a, np.ascontiguousarray(getattr(self, a)))
        return (_np_ptr(self.m_pos), _np_ptr(self.m_he), len(self.m_he), _np_ptr(self.m_t), len(self.m_t),
                _np_ptr(self.m_v), len(self.m_v), int(self.periodic))

    def delonize(self):
        """Host_Triangulation.h:305-319: Lawson flips (exact incircle) until
        every edge is locally Delaunay; False when a reflex quad (a tangled
        mesh) was left unflipped"""
        f = ctypes.c_longlong()
        rc = lib().mcl_hmesh_delonize(*self._arrays(), None, 0, ctypes.byref(f))
        self.last_flips = f.value
        return rc == 0

    def untangle(self):
        """Host_Triangulation.h:417-423: one pass of the B/C steps; True when
        something was flipped (the reference's return value)"""
        f = ctypes.c_longlong()
        check(lib().mcl_hmesh_untangle(*self._arrays(), ctypes.byref(f)), "mcl_hmesh_untangle")
        self.last_flips = f.value
        return f.value > 0

    def _move_vertex(self, p_index, d, mode):
        d = np.asarray(d, np.float32).reshape(2)
        f = ctypes.c_longlong()
        for a in ("m_pos", "m_v", "m_he", "m_t"):
            setattr(self, a, np.ascontiguousarray(getattr(self, a)))
        rc = lib().mcl_hmesh_move_vertex(_np_ptr(self.m_pos), _np_ptr(self.m_he), len(self.m_he), _np_ptr(self.m_t),
                                         len(self.m_t), _np_ptr(self.m_v), len(self.m_v), int(self.periodic),
                                         int(p_index), _np_ptr(d), int(mode), ctypes.byref(f))
        self.last_flips = f.value
        return rc == 0

    def moveFlipflop(self, p_index, d):
        """Host_Triangulation.h:77: removes vertex p_index and re-inserts it at
        its position + d (the mesh stays valid, not necessarily Delaunay)"""
        return self._move_vertex(p_index, d, _lib.MOVE_FLIPFLOP)

    def moveFlipflopDelaunay(self, p_index, d):
        """Host_Triangulation.h:78: moveFlipflop, then Lawson flips"""
        return self._move_vertex(p_index, d, _lib.MOVE_FLIPFLOP_DELAUNAY)

    def moveUntangling(self, p_index, d):
        """Host_Triangulation.h:79: moves, then untangles (False when
        inverted triangles remain)"""
        return self._move_vertex(p_index, d, _lib.MOVE_UNTANGLING)

    def moveUntanglingDelaunay(self, p_index, d):
        """Host_Triangulation.h:80: moveUntangling, then Lawson flips"""
        return self._move_vertex(p_index, d, _lib.MOVE_UNTANGLING_DELAUNAY)

    def save_state(self, path):
        """writes m_pos, m_v, m_he, m_t to a state file (mcl_state_save)"""
        check(lib().mcl_state_save(str(path).encode(), int(self.periodic),
                                   _np_ptr(np.ascontiguousarray(self.m_pos, np.float32)), len(self.m_pos),
                                   _np_ptr(np.ascontiguousarray(self.m_v, np.int32)), len(self.m_v),
                                   _np_ptr(np.ascontiguousarray(self.m_he, np.int32)), len(self.m_he),
                                   _np_ptr(np.ascontiguousarray(self.m_t, np.int32)), len(self.m_t)),
              "mcl_state_save")

    @classmethod
    def load_state(cls, path):
        """a host triangulation holding a state file's arrays (mcl_state_load);
        DeviceTriangulation(it) continues from them with fresh round
        counters (DeviceTriangulation.load_state restores those too)"""
        L = lib()
        q = [ctypes.c_int() for _ in range(5)]
        check(L.mcl_state_info(str(path).encode(), *[ctypes.byref(x) for x in q]), "mcl_state_info")
        periodic, npos, nv, nhe, nt = (x.value for x in q)
        h = cls.__new__(cls)
        h.periodic = bool(periodic)
        h.m_pos = np.zeros((npos, 2), np.float32)
        h.m_v = np.zeros((nv, 2), np.int32)
        h.m_he = np.zeros((nhe, 4), np.int32)
        h.m_t = np.zeros(nt, np.int32)
        check(L.mcl_state_load(str(path).encode(), _np_ptr(h.m_pos), _np_ptr(h.m_v), _np_ptr(h.m_he), _np_ptr(h.m_t)),
              "mcl_state_load")
        h.perm = np.zeros(0, np.int32)
        h._points = h.m_pos[4:] if not h.periodic else h.m_pos
        return h


class HostTriangulationBatch:
    """An ensemble of independent box triangulations held as one mesh.

    The reference builds one HostTriangulation per instance
    (Host_Triangulation.h:107-137, 296-300); here every instance is built on
    the host (one per thread) and the results are concatenated, so that one
    DeviceTriangulation advances the whole ensemble with each launch.
    Vertex layout: the 4 corners of every instance first (instance i:
    4i..4i+3), then each instance's points in its own Hilbert order.
    `points` is a list of (n_i, 2) arrays.  Attributes: m_pos, m_v, m_he,
    m_t (the AoS arrays of the union), vstart / hestart / tstart (ninst + 1
    offsets), perm (vertex order -> index into the concatenated input).
    """

    periodic = False

    def __init__(self, points, nthreads=0):
        pts = [np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1, 2)) for p in points]
        self.ninst = len(pts)
        counts = np.array([len(p) for p in pts], np.int32)
        allpts = np.ascontiguousarray(np.concatenate(pts) if pts else np.zeros((0, 2), np.float32))
        L = lib()
        h = _vp()
        check(L.mcl_hmesh_build_batch(_np_ptr(allpts), _np_ptr(counts), self.ninst, int(nthreads),
                                      ctypes.byref(h)), "mcl_hmesh_build_batch")
        try:
            sz = [ctypes.c_int() for _ in range(4)]
            check(L.mcl_hmesh_sizes(h, *[ctypes.byref(s) for s in sz]), "mcl_hmesh_sizes")
            npos, nv, nhe, nt = (s.value for s in sz)
            self.m_pos = np.zeros((npos, 2), np.float32)
            self.m_v = np.zeros((nv, 2), np.int32)
            self.m_he = np.zeros((nhe, 4), np.int32)
            self.m_t = np.zeros(nt, np.int32)
            self.perm = np.zeros(len(allpts), np.int32)
            check(L.mcl_hmesh_export(h, _np_ptr(self.m_pos), _np_ptr(self.m_v), _np_ptr(self.m_he),
                                     _np_ptr(self.m_t), _np_ptr(self.perm)), "mcl_hmesh_export")
            self.vstart = np.zeros(self.ninst + 1, np.int32)
            self.hestart = np.zeros(self.ninst + 1, np.int32)
            self.tstart = np.zeros(self.ninst + 1, np.int32)
            check(L.mcl_hmesh_batch_offsets(h, _np_ptr(self.vstart), _np_ptr(self.hestart), _np_ptr(self.tstart)),
                  "mcl_hmesh_batch_offsets")
        finally:
            L.mcl_hmesh_free(h)
        self.counts = counts

    def num_points(self):
        return int(self.counts.sum())

    def instance_vertices(self, i):
        """vertex ids of instance i: its 4 corners, then its points"""
        return np.concatenate([np.arange(4 * i, 4 * i + 4), np.arange(self.vstart[i], self.vstart[i + 1])])


class DeviceTriangulation:
    """Device triangulation (Device_Triangulation.h:128-223), SoA on the GPU."""

    blocksize = 256

    def __init__(self, host):
        self.h_triangulation = host
        L = lib()
        self._h = _vp()
        check(L.mcl_tri_create(_np_ptr(np.ascontiguousarray(host.m_pos)), len(host.m_pos),
                               _np_ptr(np.ascontiguousarray(host.m_v)), len(host.m_v),
                               _np_ptr(np.ascontiguousarray(host.m_he)), len(host.m_he),
                               _np_ptr(np.ascontiguousarray(host.m_t)), len(host.m_t),
                               int(host.periodic), ctypes.byref(self._h)), "mcl_tri_create")
        self.m_pos_size = len(host.m_pos)
        self.m_v_size = len(host.m_v)
        self.m_he_size = len(host.m_he)
        self.m_t_size = len(host.m_t)
        self.periodic = host.periodic
        self.last_stats = None

    @classmethod
    def from_points(cls, points, periodic=False):
        """HostTriangulation(periodic).addDelaunayPoints(points) +
        DeviceTriangulation(host) built on the GPU (mcl_tri_build), with the
        host build's vertex numbering.  points: (n, 2) float32 numpy array or
        CUDA tensor.  The returned object's h_triangulation holds the arrays
        of the result (filled by transferToHost) and perm (vertex base + k =
        points[perm[k]], base 4 in a box, 0 on the torus)."""
        import torch
        if isinstance(points, torch.Tensor):
            d = points.detach().to(device="cuda", dtype=torch.float32).contiguous().reshape(-1, 2)
        else:
            d = torch.from_numpy(np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 2))).cuda()
        n = d.shape[0]
        perm = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        self = cls.__new__(cls)
        self._h = _vp()
        torch.cuda.synchronize()
        check(lib().mcl_tri_build(_dev_ptr(d), n, int(bool(periodic)), _dev_ptr(perm), ctypes.byref(self._h), None),
              "mcl_tri_build")
        base, extra_he, extra_t = (0, 0, 0) if periodic else (4, 10, 2)
        host = HostTriangulation.__new__(HostTriangulation)
        host.periodic = bool(periodic)
        host.m_pos = np.zeros((n + base, 2), np.float32)
        host.m_v = np.zeros((n + base, 2), np.int32)
        host.m_he = np.zeros((6 * n + extra_he, 4), np.int32)
        host.m_t = np.zeros(2 * n + extra_t, np.int32)
        host.perm = perm[:n].cpu().numpy()
        host._points = np.zeros((0, 2), np.float32)
        self.h_triangulation = host
        self.m_pos_size = self.m_v_size = n + base
        self.m_he_size = 6 * n + extra_he
        self.m_t_size = 2 * n + extra_t
        self.periodic = bool(periodic)
        self.last_stats = None
        self.transferToHost()
        host._points = host.m_pos[base:].copy()
        return self

    @classmethod
    def from_ensemble(cls, points):
        """HostTriangulationBatch(points) + DeviceTriangulation(batch) built on
        the GPU (mcl_batch_build).  points: a list of (n_i, 2) float32 arrays
        (or CUDA tensors).  The result's h_triangulation is a
        HostTriangulationBatch holding the arrays of the result (filled by
        transferToHost), counts, vstart and perm as the host build's;
        hestart and tstart are None (on the GPU an instance's half-edges and
        triangles are not contiguous)."""
        import torch
        parts = []
        for p in points:
            if isinstance(p, torch.Tensor):
                parts.append(p.detach().to(device="cuda", dtype=torch.float32).reshape(-1, 2))
            else:
                parts.append(torch.from_numpy(np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1, 2))).cuda())
        ninst = len(parts)
        counts = np.array([int(p.shape[0]) for p in parts], np.int32)
        d = torch.cat(parts).contiguous() if parts else torch.zeros((0, 2), dtype=torch.float32, device="cuda")
        n = int(counts.sum())
        perm = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        self = cls.__new__(cls)
        self._h = _vp()
        torch.cuda.synchronize()
        check(lib().mcl_batch_build(_dev_ptr(d), _np_ptr(counts), ninst, _dev_ptr(perm), ctypes.byref(self._h), None),
              "mcl_batch_build")
        nv = n + 4 * ninst
        host = HostTriangulationBatch.__new__(HostTriangulationBatch)
        host.ninst = ninst
        host.counts = counts
        host.vstart = (4 * ninst + np.concatenate([[0], np.cumsum(counts)])).astype(np.int32)
        host.hestart = host.tstart = None
        host.perm = perm[:n].cpu().numpy()
        host.m_pos = np.zeros((nv, 2), np.float32)
        host.m_v = np.zeros((nv, 2), np.int32)
        host.m_he = np.zeros((6 * n + 10 * ninst, 4), np.int32)
        host.m_t = np.zeros(2 * n + 2 * ninst, np.int32)
        self.h_triangulation = host
        self.m_pos_size = self.m_v_size = nv
        self.m_he_size = 6 * n + 10 * ninst
        self.m_t_size = 2 * n + 2 * ninst
        self.periodic = False
        self.last_stats = None
        self.transferToHost()
        return self

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            try:
                lib().mcl_tri_destroy(self._h)
            except Exception:
                pass
            self._h = _vp()

    @property
    def m_pos(self):
        """device address of the float2 positions (Device_Triangulation.h:139)"""
        return lib().mcl_tri_pos_ptr(self._h)

    def download_targets(self):
        """the targets of the last step (mcl_tri_target_ptr) as a host
        (V, 2) float32 array: every vertex is at its target after a step,
        except the last step's `duplicates` (vertices sharing one target)"""
        import torch
        t = _tensor_from_ptr(lib().mcl_tri_target_ptr(self._h), (self.m_pos_size, 2), torch.float32)
        torch.cuda.synchronize()
        return t.cpu().numpy().copy()

    def pos_tensor(self):
        """the device positions as a torch view (no copy)"""
        import torch
        return _tensor_from_ptr(self.m_pos, (self.m_pos_size, 2), torch.float32)

    # ------------------------------------------------------------ transfers
    def save_state(self, path):
        """the current triangulation to a state file (mcl_tri_save);
        DeviceTriangulation.load_state(path) resumes it, and
        HostTriangulation.load_state(path) reads its arrays back"""
        check(lib().mcl_tri_save(self._h, str(path).encode()), "mcl_tri_save")

    @classmethod
    def load_state(cls, path):
        """a device triangulation resumed from a state file (mcl_tri_load):
        the arrays and the flip rounds' counters of the saved object, so that
        the same steps give the same positions as they would have given it.
        h_triangulation holds the file's arrays."""
        host = HostTriangulation.load_state(path)
        self = cls.__new__(cls)
        self._h = _vp()
        check(lib().mcl_tri_load(str(path).encode(), ctypes.byref(self._h)), "mcl_tri_load")
        self.h_triangulation = host
        self.m_pos_size = self.m_v_size = len(host.m_pos)
        self.m_he_size = len(host.m_he)
        self.m_t_size = len(host.m_t)
        self.periodic = host.periodic
        self.last_stats = None
        return self

    def transferToHost(self):
        h = self.h_triangulation
        check(lib().mcl_tri_download(self._h, _np_ptr(h.m_pos), _np_ptr(h.m_v), _np_ptr(h.m_he),
                                     _np_ptr(h.m_t)), "mcl_tri_download")
        return True

    def transferToDevice(self):
        h = self.h_triangulation
        check(lib().mcl_tri_upload(self._h, _np_ptr(np.ascontiguousarray(h.m_pos)),
                                   _np_ptr(np.ascontiguousarray(h.m_v)), _np_ptr(np.ascontiguousarray(h.m_he)),
                                   _np_ptr(np.ascontiguousarray(h.m_t))), "mcl_tri_upload")
        return True

    def download_positions(self):
        pos = np.zeros((self.m_pos_size, 2), np.float32)
        check(lib().mcl_tri_download(self._h, _np_ptr(pos), None, None, None), "mcl_tri_download")
        return pos

    def download(self):
        pos = np.zeros((self.m_pos_size, 2), np.float32)
        v = np.zeros((self.m_v_size, 2), np.int32)
        he = np.zeros((self.m_he_size, 4), np.int32)
        t = np.zeros(self.m_t_size, np.int32)
        check(lib().mcl_tri_download(self._h, _np_ptr(pos), _np_ptr(v), _np_ptr(he), _np_ptr(t)),
              "mcl_tri_download")
        return pos, v, he, t

    # ---------------------------------------------------------- operations
    def _stats_call(self, fn, *args, allow=()):
        st = _lib.Stats()
        rc = fn(self._h, *args, ctypes.byref(st), None)
        self.last_stats = st.as_dict()
        if rc not in allow:
            check(rc, fn.__name__)
        return rc

    def delonize(self, max_rounds=0):
        """Lawson rounds with one launch pair and a host read per round (the
        reference's loop shape, Device_Triangulation.h:353-390)."""
        self._stats_call(lib().mcl_tri_delonize, 1, int(max_rounds))
        return True

    def delonize2(self, max_rounds=0):
        """Lawson rounds driven on the device (Device_Triangulation.h:324-351)."""
        self._stats_call(lib().mcl_tri_delonize, 0, int(max_rounds))
        return True

    def untangle(self, max_rounds=0):
        """B/C-step rounds (Device_Triangulation.h:422-459); False if
        inverted triangles remain."""
        rc = self._stats_call(lib().mcl_tri_untangle, int(max_rounds), allow=(0, 6))
        return rc == 0

    untangle2 = untangle  # Device_Triangulation.h:461-495

    def get_option(self, option):
        v = ctypes.c_longlong()
        check(lib().mcl_tri_get_option(self._h, int(option), ctypes.byref(v)), "mcl_tri_get_option")
        return v.value

    def options(self):
        """every tunable option by its tuning-table name"""
        return {name: self.get_option(o) for name, o in _lib.TRI_OPTIONS.items()}

    def move_vertices(self, ids, d, mode):
        """mcl_tri_move_vertices: ids (host ints) move by d (host (n, 2)
        float32) with mode MOVE_*; returns the status (0 when every listed
        vertex is at its target and, in the untangling modes, nothing stays
        inverted)"""
        ids = np.ascontiguousarray(np.asarray(ids, np.int32).reshape(-1))
        d = np.ascontiguousarray(np.asarray(d, np.float32).reshape(-1, 2))
        st = _lib.Stats()
        rc = lib().mcl_tri_move_vertices(self._h, _np_ptr(ids), _np_ptr(d), len(ids), int(mode), ctypes.byref(st),
                                         None)
        self.last_stats = st.as_dict()
        return rc

    def moveFlipflop(self, p_index, d):
        """Device_Triangulation.h:183 (a stub in the reference): remove and
        re-insert one vertex at its position + d"""
        return self.move_vertices([p_index], [d], _lib.MOVE_FLIPFLOP) == 0

    def moveFlipflopDelaunay(self, p_index, d):
        """Device_Triangulation.h:184: moveFlipflop, then Lawson rounds"""
        return self.move_vertices([p_index], [d], _lib.MOVE_FLIPFLOP_DELAUNAY) == 0

    def moveUntangling(self, p_index, d):
        """Device_Triangulation.h:185: move, then untangle rounds"""
        return self.move_vertices([p_index], [d], _lib.MOVE_UNTANGLING) == 0

    def moveUntanglingDelaunay(self, p_index, d):
        """Device_Triangulation.h:186: moveUntangling, then Lawson rounds"""
        return self.move_vertices([p_index], [d], _lib.MOVE_UNTANGLING_DELAUNAY) == 0

    def movePoints(self, d):
        """Device_Triangulation.h:395-402: d is a device (V,2) float32 tensor"""
        check(lib().mcl_tri_move(self._h, _dev_ptr(d), None), "mcl_tri_move")
        return True

    def integrateMovePoints(self, dt, v):
        """Device_Triangulation.h:403-410"""
        check(lib().mcl_tri_integrate(self._h, _dev_ptr(dt), _dev_ptr(v), None), "mcl_tri_integrate")
        return True

    def step(self, motion, sigma, amp=0.0, seed=1, step=0, variant=0, max_iterations=0):
        """one kinetic step (move, keep valid, restore Delaunay)"""
        p = _lib.StepParams(int(motion), float(sigma), float(amp), int(seed), int(step), int(variant),
                            int(max_iterations))
        st = _lib.Stats()
        rc = lib().mcl_tri_step(self._h, ctypes.byref(p), ctypes.byref(st), None)
        self.last_stats = st.as_dict()
        check(rc, "mcl_tri_step")
        return self.last_stats

    def timestep(self, motion, sigma, amp=0.0, seed=1, step=0, variant=0, ring=None, grid=None, rad=0.0,
                 frnn=None, bgrid=None):
        """one whole timestep (mcl_tri_timestep): the kinetic step, then
        asynchronously the CSR one-ring into ring = (offsets, nbrs) and the
        grid's sort of the new positions with its CSR FRNN query into
        frnn = (offsets, nbrs) (grid: a GridCount2d over every vertex; or,
        for an ensemble, bgrid: a GridCount2dBatch, mcl_tri_timestep_batch)"""
        p = _lib.StepParams(int(motion), float(sigma), float(amp), int(seed), int(step), int(variant), 0)
        st = _lib.Stats()
        ro, rn = ring if ring is not None else (None, None)
        fo, fn = frnn if frnn is not None else (None, None)
        if bgrid is not None:
            if grid is not None:
                raise ValueError("grid or bgrid, not both")
            rc = lib().mcl_tri_timestep_batch(self._h, ctypes.byref(p), _dev_ptr(ro), _dev_ptr(rn),
                                              0 if rn is None else int(rn.numel()), bgrid._g, _dev_ptr(fo),
                                              _dev_ptr(fn), 0 if fn is None else int(fn.numel()),
                                              ctypes.byref(st), None)
            self.last_stats = st.as_dict()
            check(rc, "mcl_tri_timestep_batch")
            return self.last_stats
        rc = lib().mcl_tri_timestep(self._h, ctypes.byref(p), _dev_ptr(ro), _dev_ptr(rn),
                                    0 if rn is None else int(rn.numel()), None if grid is None else grid._g,
                                    float(rad), _dev_ptr(fo), _dev_ptr(fn), 0 if fn is None else int(fn.numel()),
                                    ctypes.byref(st), None)
        self.last_stats = st.as_dict()
        check(rc, "mcl_tri_timestep")
        return self.last_stats

    def move_legalize(self, d, variant=0):
        """movePoints(d) + untangle + delonize as one call that always ends
        valid and exactly Delaunay (d: device (V,2) float32 tensor)"""
        st = _lib.Stats()
        rc = lib().mcl_tri_move_legalize(self._h, _dev_ptr(d), int(variant), ctypes.byref(st), None)
        self.last_stats = st.as_dict()
        check(rc, "mcl_tri_move_legalize")
        return self.last_stats

    def count_inverted(self):
        out = ctypes.c_longlong()
        check(lib().mcl_tri_count_inverted(self._h, ctypes.byref(out), None), "mcl_tri_count_inverted")
        return out.value

    def count_non_delaunay(self):
        out = ctypes.c_longlong()
        check(lib().mcl_tri_count_non_delaunay(self._h, ctypes.byref(out), None), "mcl_tri_count_non_delaunay")
        return out.value

    # ------------------------------------------------------------- queries
    def _one_ring(self, K, layout, ring):
        ov = ctypes.c_int()
        rc = lib().mcl_tri_one_ring(self._h, int(K), layout, _dev_ptr(ring), ctypes.byref(ov), None)
        self.ring_overflow = ov.value
        if rc != _lib.MCL_ERR_CAPACITY:
            check(rc, "mcl_tri_one_ring")
        return rc == 0

    def oneRing(self, K, ring):
        """oneRing<K>(int*): SoA ring[(c+1)*V + i], count in ring[i]
        (kernels.cuh:54-84).  False (ring_overflow = the number of vertices)
        when a vertex has more than K - 1 neighbours; those keep the first K - 1"""
        return self._one_ring(K, 0, ring)

    def oneRing2(self, K, ring):
        """oneRing2<K>(int*): AoS ring[(c+1) + K*i], count in ring[K*i]
        (kernels.cuh:91-121); overflow as oneRing"""
        return self._one_ring(K, 1, ring)

    def one_ring_csr(self, offsets, nbrs=None):
        """CSR one-ring (no cap); returns the entry count"""
        tot = ctypes.c_int()
        cap = 0 if nbrs is None else int(nbrs.numel())
        rc = lib().mcl_tri_one_ring_csr(self._h, _dev_ptr(offsets), _dev_ptr(nbrs), cap, ctypes.byref(tot), None)
        if nbrs is not None:
            check(rc, "mcl_tri_one_ring_csr")
        return tot.value

    def closestNeighbors(self, offsets, nbrs, out):
        check(lib().mcl_tri_closest(self._h, _dev_ptr(offsets), _dev_ptr(nbrs), _dev_ptr(out), None),
              "mcl_tri_closest")
        return True

    def closestNeighbors_ring(self, K, ring, out, layout=0):
        """closestNeighbors<K>(ring, out) (layout 0) / closestNeighbors2<K> (layout 1)
        over the ring written by oneRing / oneRing2"""
        check(lib().mcl_tri_closest_ring(self._h, int(K), int(layout), _dev_ptr(ring), _dev_ptr(out), None),
              "mcl_tri_closest_ring")
        return True

    def getFRNN(self, K, Mx, r, ring, out, layout=0):
        """getFRNN<K, M>(r, ring, out) (layout 0) / getFRNN2<K, M> (layout 1)"""
        ov = ctypes.c_int()
        check(lib().mcl_tri_frnn_graph(self._h, float(r), int(K), int(layout), _dev_ptr(ring), int(Mx), _dev_ptr(out),
                                       ctypes.byref(ov), None), "mcl_tri_frnn_graph")
        return True

    def one_ring_csr_async(self, offsets, nbrs):
        """CSR one-ring without a host read-back (nbrs must hold m_he_size ints)"""
        check(lib().mcl_tri_one_ring_csr(self._h, _dev_ptr(offsets), _dev_ptr(nbrs), int(nbrs.numel()), None, None),
              "mcl_tri_one_ring_csr")

    def set_option(self, option, value):
        """mcl_tri_set_option: OPT_STAR_MIN_VERTICES / OPT_STAR_ENABLE (fast step choices)"""
        check(lib().mcl_tri_set_option(self._h, int(option), int(value)), "mcl_tri_set_option")

    def kernel_timer(self, enable=True):
        check(lib().mcl_tri_kernel_timer(self._h, int(enable)), "mcl_tri_kernel_timer")

    def kernel_timer_read(self, which):
        """(total GPU ms, launches) of kernel class `which` (0 Lawson full round, 1 check all, 2 move,
        3 relocation apply, 4 take-back apply)"""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        check(lib().mcl_tri_kernel_timer_read(self._h, int(which), ctypes.byref(ms), ctypes.byref(n)),
              "mcl_tri_kernel_timer_read")
        return ms.value, n.value


def _tensor_from_ptr(ptr, shape, dtype):
    """wrap a device allocation owned by the library as a torch tensor"""
    import torch

    class _Holder:
        def __init__(self, p, nbytes):
            self.__cuda_array_interface__ = {
                "shape": shape, "typestr": "<f4" if dtype == torch.float32 else "<i4",
                "data": (p, False), "version": 2}

    n = int(np.prod(shape))
    return torch.as_tensor(_Holder(ptr, n * 4), device="cuda")


class GridCount2d:
    """Uniform-grid FRNN (gridCount2d.h:23-48)."""

    blocksize = 256

    def __init__(self, numP, min_xy, cell_size, num_cells, flags=_lib.GRID_ROWMAJOR):
        self.numP = int(numP)
        self._g = _vp()
        check(lib().mcl_grid_create(self.numP, float(min_xy[0]), float(min_xy[1]), float(cell_size[0]),
                                    float(cell_size[1]), int(num_cells[0]), int(num_cells[1]), int(flags),
                                    ctypes.byref(self._g)), "mcl_grid_create")

    def __del__(self):
        if getattr(self, "_g", None) and self._g.value:
            try:
                lib().mcl_grid_destroy(self._g)
            except Exception:
                pass
            self._g = _vp()

    def update(self, pos, vel=None):
        """sorts the device arrays in place by cell (gridCount2d.h:119-135)"""
        check(lib().mcl_grid_update(self._g, _dev_ptr(pos), _dev_ptr(vel), None), "mcl_grid_update")

    def index_ptrs(self):
        ind, perm, start, cnt = _vp(), _vp(), _vp(), _vp()
        check(lib().mcl_grid_index_ptrs(self._g, ctypes.byref(ind), ctypes.byref(perm), ctypes.byref(start),
                                        ctypes.byref(cnt)), "mcl_grid_index_ptrs")
        return ind.value, perm.value, start.value

    def d_ind(self):
        """new slot of each previous index (gridCount2d.h:40) as a torch int32 tensor"""
        import torch
        ind, _, _ = self.index_ptrs()
        return _tensor_from_ptr(ind, (self.numP,), torch.int32).clone()

    def perm(self):
        import torch
        _, perm, _ = self.index_ptrs()
        return _tensor_from_ptr(perm, (self.numP,), torch.int32).clone()

    def frnn_csr(self, pos, rad, offsets, nbrs=None):
        """pairs within rad of each sorted slot, as CSR over slots; returns the pair count"""
        tot = ctypes.c_longlong()
        cap = 0 if nbrs is None else int(nbrs.numel())
        rc = lib().mcl_grid_frnn_csr(self._g, _dev_ptr(pos), float(rad), _dev_ptr(offsets), _dev_ptr(nbrs), cap,
                                     ctypes.byref(tot), None)
        if nbrs is not None:
            check(rc, "mcl_grid_frnn_csr")
        return tot.value

    def apply_f_frnn(self, f, pos, rad):
        """calls f(i, j, d, dist) with vectors of all pairs (i, j) within rad,
        i == j included (gridCount2d.h:262-302)"""
        import torch
        n = self.numP
        off = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
        tot = self.frnn_csr(pos, rad, off)
        nbr = torch.zeros(max(tot, 1), dtype=torch.int32, device="cuda")
        self.frnn_csr(pos, rad, off, nbr)
        counts = (off[1:] - off[:-1]).long()
        i = torch.repeat_interleave(torch.arange(n, device="cuda"), counts)
        j = nbr[:tot].long()
        i = torch.cat([i, torch.arange(n, device="cuda")])
        j = torch.cat([j, torch.arange(n, device="cuda")])
        d = pos[j] - pos[i]
        f(i, j, d, torch.sqrt((d.double() ** 2).sum(1)))

    def set_block(self, block):
        """threads per block of this grid's sort and query kernels (64, 128, 256)"""
        check(lib().mcl_grid_set_option(self._g, _lib.GRID_OPT_BLOCK, int(block)), "mcl_grid_set_option")

    def block(self):
        v = ctypes.c_longlong()
        check(lib().mcl_grid_get_option(self._g, _lib.GRID_OPT_BLOCK, ctypes.byref(v)), "mcl_grid_get_option")
        return v.value

    def build(self, src_pos):
        """counting sort of src_pos into the grid's own sorted copy (src untouched, async)"""
        check(lib().mcl_grid_build(self._g, _dev_ptr(src_pos), None), "mcl_grid_build")

    def sorted_ptr(self):
        return lib().mcl_grid_sorted_ptr(self._g)

    def frnn_csr_async(self, pos_ptr, rad, offsets, nbrs):
        """CSR FRNN over sorted slots without a host read-back"""
        check(lib().mcl_grid_frnn_csr(self._g, _vp(pos_ptr) if isinstance(pos_ptr, int) else _dev_ptr(pos_ptr),
                                      float(rad), _dev_ptr(offsets), _dev_ptr(nbrs), int(nbrs.numel()), None, None),
              "mcl_grid_frnn_csr")

    def kernel_timer(self, enable=True):
        check(lib().mcl_grid_kernel_timer(self._g, int(enable)), "mcl_grid_kernel_timer")

    def kernel_timer_read(self, which):
        """(total GPU ms, launches) of kernel class `which` (0 FRNN count, 1 FRNN fill, 2 build)"""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        check(lib().mcl_grid_kernel_timer_read(self._g, int(which), ctypes.byref(ms), ctypes.byref(n)),
              "mcl_grid_kernel_timer_read")
        return ms.value, n.value

    def mean_num_particle_in_cell(self):
        out = ctypes.c_float()
        check(lib().mcl_grid_mean_occupancy(self._g, ctypes.byref(out), None), "mcl_grid_mean_occupancy")
        return out.value


class GridCount2dBatch:
    """One uniform grid per instance of an ensemble, sorted and queried in
    single launches (each instance of the reference would hold its own
    GridCount2d, gridCount2d.h:23-48).  vstart: the ensemble's ninst + 1
    point offsets (HostTriangulationBatch.vstart); radius: one query radius
    per instance (cells of max(3, floor(1/r_i)) per side)."""

    def __init__(self, vstart, radius):
        vstart = np.ascontiguousarray(vstart, np.int32)
        radius = np.ascontiguousarray(radius, np.float32)
        self.ninst = len(vstart) - 1
        if len(radius) != self.ninst:
            raise ValueError("one radius per instance")
        self._g = _vp()
        check(lib().mcl_bgrid_create(self.ninst, _np_ptr(vstart), _np_ptr(radius), ctypes.byref(self._g)),
              "mcl_bgrid_create")
        n, first, cells = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().mcl_bgrid_sizes(self._g, ctypes.byref(n), ctypes.byref(first), ctypes.byref(cells)),
              "mcl_bgrid_sizes")
        self.numP, self.first, self.ncells = n.value, first.value, cells.value

    def __del__(self):
        if getattr(self, "_g", None) and self._g.value:
            try:
                lib().mcl_bgrid_destroy(self._g)
            except Exception:
                pass
            self._g = _vp()

    def build(self, pos):
        """counting sort of every instance's points (pos: the ensemble's device positions; async)"""
        check(lib().mcl_bgrid_build(self._g, _dev_ptr(pos), None), "mcl_bgrid_build")

    def sorted_ptr(self):
        return lib().mcl_bgrid_sorted_ptr(self._g)

    def sorted_pos(self):
        import torch
        return _tensor_from_ptr(self.sorted_ptr(), (self.numP, 2), torch.float32)

    def index_ptrs(self):
        ind, perm, start = _vp(), _vp(), _vp()
        check(lib().mcl_bgrid_index_ptrs(self._g, ctypes.byref(ind), ctypes.byref(perm), ctypes.byref(start)),
              "mcl_bgrid_index_ptrs")
        return ind.value, perm.value, start.value

    def perm(self):
        """point (vertex id - first) of each sorted slot, a torch int32 copy"""
        import torch
        _, perm, _ = self.index_ptrs()
        return _tensor_from_ptr(perm, (self.numP,), torch.int32).clone()

    def frnn_csr(self, offsets, nbrs=None):
        """pairs within each instance's radius, CSR over sorted slots; returns the pair count"""
        tot = ctypes.c_longlong()
        cap = 0 if nbrs is None else int(nbrs.numel())
        rc = lib().mcl_bgrid_frnn_csr(self._g, _dev_ptr(offsets), _dev_ptr(nbrs), cap, ctypes.byref(tot), None)
        if nbrs is not None:
            check(rc, "mcl_bgrid_frnn_csr")
        return tot.value

    def frnn_csr_async(self, offsets, nbrs):
        check(lib().mcl_bgrid_frnn_csr(self._g, _dev_ptr(offsets), _dev_ptr(nbrs), int(nbrs.numel()), None, None),
              "mcl_bgrid_frnn_csr")

    def kernel_timer(self, enable=True):
        check(lib().mcl_bgrid_kernel_timer(self._g, int(enable)), "mcl_bgrid_kernel_timer")

    def kernel_timer_read(self, which):
        """(total GPU ms, launches) of kernel class `which` (0 FRNN count, 1 FRNN fill, 2 build)"""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        check(lib().mcl_bgrid_kernel_timer_read(self._g, int(which), ctypes.byref(ms), ctypes.byref(n)),
              "mcl_bgrid_kernel_timer_read")
        return ms.value, n.value
