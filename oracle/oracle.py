"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f = ctypes.c_float
_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_ll = ctypes.c_longlong


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        sig = {
            "orc_orient2d_ref": (_i, [_f] * 6),
            "orc_incircle_ref": (_i, [_f] * 8),
            "orc_orient2d_exact": (_i, [_p, _p, _p, _i]),
            "orc_incircle_exact": (_i, [_p, _p, _p, _p, _i]),
            "orc_incircle_sos": (_i, [_p, _p, _p, _p, _i]),
            "orc_pred_stats": (None, [_p]),
            "orc_pred_stats_reset": (None, []),
            "orc_tri_create": (_p, [_i]),
            "orc_tri_destroy": (None, [_p]),
            "orc_tri_sizes": (_i, [_p, _p]),
            "orc_tri_add_delaunay_points": (_i, [_p, _p, _i, _i]),
            "orc_tri_export": (None, [_p, _p, _p, _p, _p]),
            "orc_tri_import": (_i, [_p, _p, _i, _p, _i, _p, _i, _p, _i]),
            "orc_tri_set_positions": (None, [_p, _p, _i]),
            "orc_tri_delonize": (_i, [_p, _p]),
            "orc_tri_delonize_worklist": (_i, [_p, _p]),
            "orc_tri_untangle": (_i, [_p]),
            "orc_tri_move_point": (_i, [_p, _i, _f, _f]),
            "orc_tri_set_periodic": (None, [_p, _i]),
            "orc_tri_diagnostics": (None, [_p, _p]),
            "orc_tri_one_ring": (_i, [_p, _i, _p, _i]),
            "orc_tri_get_frnn": (_i, [_p, _i, _f, _p, _i]),
            "orc_check_mesh": (_i, [_p, _i, _p, _i, _p, _i, _p, _i, _i, _i, ctypes.c_char_p, _i]),
            "orc_count_non_delaunay": (_ll, [_p, _p, _i, _i]),
            "orc_canonical_triangles": (_i, [_p, _i, _p, _i, _p]),
            "orc_torus_delaunay": (_i, [_p, _i, _f, _p]),
            "orc_grid_frnn": (_i, [_p, _i, _f, _f, _f, _f, _i, _i, _f, _i, _p, _p, _i]),
            "orc_brute_frnn": (_i, [_p, _i, _f, _i, _p, _p, _i]),
            "orc_grid_sort": (_i, [_p, _i, _f, _f, _f, _f, _i, _i, _p, _p, _p]),
            "orc_mt_uniform_floats": (None, [ctypes.c_uint32, _f, _f, _i, _p]),
            "orc_mt_uniform_doubles": (None, [ctypes.c_uint32, _d, _d, _i, _p]),
            "orc_mt_raw": (None, [ctypes.c_uint32, _i, _p]),
            "orc_exp_product": (_i, [_i, _p, _i, _p, _p]),
            "orc_expansion_sign": (_i, [_i, _p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


# ------------------------------------------------------------- predicates
def orient_ref(a, b, c):
    return lib().orc_orient2d_ref(*map(float, (*a, *b, *c)))


def incircle_ref(a, b, c, d):
    return lib().orc_incircle_ref(*map(float, (*a, *b, *c, *d)))


def _dv(p):
    return np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(2))


def orient_exact(a, b, c, periodic=False):
    a, b, c = _dv(a), _dv(b), _dv(c)
    return lib().orc_orient2d_exact(_ptr(a), _ptr(b), _ptr(c), int(periodic))


def incircle_exact(a, b, c, d, periodic=False):
    a, b, c, d = _dv(a), _dv(b), _dv(c), _dv(d)
    return lib().orc_incircle_exact(_ptr(a), _ptr(b), _ptr(c), _ptr(d), int(periodic))


def incircle_sos(a, b, c, d, periodic=False):
    a, b, c, d = _dv(a), _dv(b), _dv(c), _dv(d)
    return lib().orc_incircle_sos(_ptr(a), _ptr(b), _ptr(c), _ptr(d), int(periodic))


def pred_stats(reset=False):
    out = np.zeros(4, dtype=np.int64)
    lib().orc_pred_stats(_ptr(out))
    if reset:
        lib().orc_pred_stats_reset()
    return out


# ----------------------------------------------------------- triangulation
class HostTri:
    """Restated HostTriangulation (Host_Triangulation.h:24-98)."""

    REFERENCE = 0
    EXACT = 1

    def __init__(self, mode=EXACT):
        self.h = lib().orc_tri_create(mode)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tri_destroy(self.h)
            self.h = None

    def sizes(self):
        s = np.zeros(4, dtype=np.int32)
        lib().orc_tri_sizes(self.h, _ptr(s))
        return tuple(int(x) for x in s)

    def add_delaunay_points(self, xy, sort_mode=0):
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
        r = lib().orc_tri_add_delaunay_points(self.h, _ptr(xy), len(xy), sort_mode)
        if r < 0:
            raise ValueError("point %d could not be inserted" % (-r - 1))
        return r

    def export(self):
        npos, nv, nhe, nt = self.sizes()
        pos = np.zeros((npos, 2), np.float32)
        v = np.zeros((nv, 2), np.int32)
        he = np.zeros((nhe, 4), np.int32)
        t = np.zeros(nt, np.int32)
        lib().orc_tri_export(self.h, _ptr(pos), _ptr(v), _ptr(he), _ptr(t))
        return pos, v, he, t

    def load(self, pos, v, he, t):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
        v = np.ascontiguousarray(v, np.int32).reshape(-1, 2)
        he = np.ascontiguousarray(he, np.int32).reshape(-1, 4)
        t = np.ascontiguousarray(t, np.int32).reshape(-1)
        lib().orc_tri_import(self.h, _ptr(pos), len(pos), _ptr(v), len(v), _ptr(he), len(he),
                             _ptr(t), len(t))

    def set_periodic(self, periodic=True):
        lib().orc_tri_set_periodic(self.h, int(periodic))

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
        lib().orc_tri_set_positions(self.h, _ptr(pos), len(pos))

    def delonize(self):
        f = ctypes.c_longlong(0)
        lib().orc_tri_delonize(self.h, ctypes.byref(f))
        return f.value

    def delonize_worklist(self):
        f = ctypes.c_longlong(0)
        lib().orc_tri_delonize_worklist(self.h, ctypes.byref(f))
        return f.value

    def untangle(self):
        return lib().orc_tri_untangle(self.h)

    def move_point(self, i, d):
        return lib().orc_tri_move_point(self.h, int(i), float(d[0]), float(d[1]))

    def diagnostics(self):
        out = np.zeros(4, np.int64)
        lib().orc_tri_diagnostics(self.h, _ptr(out))
        return dict(non_delaunay=int(out[0]), creased=int(out[1]), inverted=int(out[2]),
                    non_ccw=int(out[3]))

    def one_ring(self, v, cap=256):
        out = np.zeros(cap, np.int32)
        n = lib().orc_tri_one_ring(self.h, int(v), _ptr(out), cap)
        return out[: min(n, cap)].tolist()

    def get_frnn(self, v, r, cap=4096):
        out = np.zeros(cap, np.int32)
        n = lib().orc_tri_get_frnn(self.h, int(v), float(r), _ptr(out), cap)
        return out[: min(n, cap)].tolist()


def check_mesh(pos, v, he, t, periodic=False, require_ccw=True):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
    v = np.ascontiguousarray(v, np.int32).reshape(-1, 2)
    he = np.ascontiguousarray(he, np.int32).reshape(-1, 4)
    t = np.ascontiguousarray(t, np.int32).reshape(-1)
    msg = ctypes.create_string_buffer(256)
    rc = lib().orc_check_mesh(_ptr(pos), len(pos), _ptr(he), len(he), _ptr(t), len(t), _ptr(v),
                              len(v), int(periodic), int(require_ccw), msg, 256)
    return rc, msg.value.decode()


def count_non_delaunay(pos, he, periodic=False):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
    he = np.ascontiguousarray(he, np.int32).reshape(-1, 4)
    return int(lib().orc_count_non_delaunay(_ptr(pos), _ptr(he), len(he), int(periodic)))


def canonical_triangles(he, t):
    he = np.ascontiguousarray(he, np.int32).reshape(-1, 4)
    t = np.ascontiguousarray(t, np.int32).reshape(-1)
    out = np.zeros((len(t), 3), np.int32)
    lib().orc_canonical_triangles(_ptr(he), len(he), _ptr(t), len(t), _ptr(out))
    return out


def canonical_from_list(tri3):
    """Canonical form of a plain (T,3) CCW triangle list."""
    tri3 = np.asarray(tri3, np.int64).reshape(-1, 3)
    r = tri3.copy()
    for _ in range(2):
        m = (r[:, 0] > r[:, 1]) | (r[:, 0] > r[:, 2])
        r[m] = np.roll(r[m], -1, axis=1)
    order = np.lexsort((r[:, 2], r[:, 1], r[:, 0]))
    return r[order].astype(np.int32)


def torus_delaunay(xy, margin=None):
    """Periodic DT on the unit torus as a (2n, 3) CCW triangle list.  The copy
    margin grows until every directed edge has its reverse (a closed torus)."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    n = len(xy)
    m = margin if margin is not None else max(0.02, 6.0 / np.sqrt(max(n, 1)))
    while True:
        out = np.zeros((2 * n, 3), np.int32)
        k = lib().orc_torus_delaunay(_ptr(xy), n, float(min(m, 0.5)), _ptr(out))
        if k == 2 * n:
            e = set()
            for a, b, c in out.tolist():
                e.update(((a, b), (b, c), (c, a)))
            if all((v, u) in e for (u, v) in e):
                return out
        if m >= 0.5:
            raise RuntimeError("torus delaunay failed (n=%d, got %d triangles)" % (n, k))
        m *= 2


def delaunay(xy, mode=HostTri.EXACT):
    """From-scratch DT of xy (vertex id = 4 + index) with the 4 reference corners."""
    T = HostTri(mode)
    T.add_delaunay_points(xy, 0)
    return T


# ------------------------------------------------------------------ grid
def grid_frnn(xy, minxy, cs, nc, rad, periodic=False, cap=256):
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    n = len(xy)
    counts = np.zeros(n, np.int32)
    nbr = np.full((n, cap), -1, np.int32)
    lib().orc_grid_frnn(_ptr(xy), n, float(minxy[0]), float(minxy[1]), float(cs[0]), float(cs[1]),
                        int(nc[0]), int(nc[1]), float(rad), int(periodic), _ptr(counts), _ptr(nbr),
                        cap)
    return counts, nbr


def brute_frnn(xy, rad, periodic=False, cap=256):
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    n = len(xy)
    counts = np.zeros(n, np.int32)
    nbr = np.full((n, cap), -1, np.int32)
    lib().orc_brute_frnn(_ptr(xy), n, float(rad), int(periodic), _ptr(counts), _ptr(nbr), cap)
    return counts, nbr


def grid_sort(xy, minxy, cs, nc):
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    n = len(xy)
    h = np.zeros(n, np.int32)
    perm = np.zeros(n, np.int32)
    start = np.zeros(nc[0] * nc[1] + 1, np.int32)
    lib().orc_grid_sort(_ptr(xy), n, float(minxy[0]), float(minxy[1]), float(cs[0]), float(cs[1]),
                        int(nc[0]), int(nc[1]), _ptr(h), _ptr(perm), _ptr(start))
    return h, perm, start


# ------------------------------------------------------------------- rng
def mt_uniform_floats(seed, a, b, n):
    out = np.zeros(n, np.float32)
    lib().orc_mt_uniform_floats(seed, a, b, n, _ptr(out))
    return out


def mt_uniform_doubles(seed, a, b, n):
    out = np.zeros(n, np.float64)
    lib().orc_mt_uniform_doubles(seed, a, b, n, _ptr(out))
    return out


def mt_raw(seed, n):
    out = np.zeros(n, np.uint32)
    lib().orc_mt_raw(seed, n, _ptr(out))
    return out
