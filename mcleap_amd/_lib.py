"""ctypes binding of libmcleap_hip.so (the C ABI declared in include/mcleap_hip.h).

The product path has no CPU fallback: importing a GPU entry point without the
built library, or calling one without a GPU, raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcleap_hip.so")

MCL_OK = 0
MCL_ERR_CAPACITY = 5
ERRORS = {
    1: "HIP runtime error",
    2: "invalid input",
    3: "duplicate point",
    4: "topology error",
    5: "capacity exceeded",
    6: "untangling left inverted triangles",
    7: "flip round cap reached",
    8: "no usable GPU",
}

MOTION_GAUSS = 0
MOTION_SHEAR = 1
GRID_ROWMAJOR = 0
GRID_MORTON = 1
GRID_PERIODIC = 2
MOVE_FLIPFLOP = 0
MOVE_FLIPFLOP_DELAUNAY = 1
MOVE_UNTANGLING = 2
MOVE_UNTANGLING_DELAUNAY = 3
OPT_STAR_MIN_VERTICES = 1
OPT_STAR_ENABLE = 2
OPT_STAR_GIVE_UP = 3
OPT_GRAPH = 4
OPT_TAIL_THRESHOLD = 5
OPT_REVERT_TAIL = 6
OPT_RELOC_TAIL = 7
OPT_BIG_GRID_MIN = 8
OPT_STAR_HEAD_PASSES = 9
GRID_OPT_BLOCK = 1
# the tuning table's parameter names and their option ids
TRI_OPTIONS = {"star_min_vertices": 1, "star_enable": 2, "star_give_up": 3, "graph": 4, "tail_threshold": 5,
               "revert_tail": 6, "reloc_tail": 7, "big_grid_min": 8, "star_head_passes": 9}


class MclError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s failed: %s (status %d)" % (what, ERRORS.get(code, "unknown"), code))
        self.code = code


class Stats(ctypes.Structure):
    """mcl_stats_t"""
    _fields_ = [
        ("rounds", ctypes.c_longlong),
        ("flips", ctypes.c_longlong),
        ("untangle_flips", ctypes.c_longlong),
        ("reverts", ctypes.c_longlong),
        ("iterations", ctypes.c_longlong),
        ("stalled", ctypes.c_longlong),
        ("inverted", ctypes.c_longlong),
        ("relocated", ctypes.c_longlong),
        ("pred_orient", ctypes.c_longlong * 4),
        ("pred_incircle", ctypes.c_longlong * 4),
        ("ms", ctypes.c_double),
        ("relocated_grid", ctypes.c_longlong),
        ("star_retries", ctypes.c_longlong),
        ("fallbacks", ctypes.c_longlong),
        ("duplicates", ctypes.c_longlong),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["pred_orient"] = list(self.pred_orient)
        d["pred_incircle"] = list(self.pred_incircle)
        return d


class StepParams(ctypes.Structure):
    """mcl_step_params_t"""
    _fields_ = [
        ("motion", ctypes.c_int),
        ("sigma", ctypes.c_float),
        ("amp", ctypes.c_float),
        ("seed", ctypes.c_ulonglong),
        ("step", ctypes.c_ulonglong),
        ("variant", ctypes.c_int),
        ("max_iterations", ctypes.c_int),
    ]


_L = None
_p = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_pp = ctypes.POINTER(ctypes.c_void_p)

SIGNATURES = {
    "mcl_hmesh_build": [_p, _i, _i, _pp],
    "mcl_hmesh_sizes": [_p, _p, _p, _p, _p],
    "mcl_hmesh_export": [_p, _p, _p, _p, _p, _p],
    "mcl_hmesh_free": [_p],
    "mcl_hmesh_build_batch": [_p, _p, _i, _i, _pp],
    "mcl_hmesh_batch_offsets": [_p, _p, _p, _p],
    "mcl_tri_create": [_p, _i, _p, _i, _p, _i, _p, _i, _i, _pp],
    "mcl_tri_destroy": [_p],
    "mcl_tri_sizes": [_p, _p, _p, _p, _p],
    "mcl_tri_download": [_p, _p, _p, _p, _p],
    "mcl_tri_upload": [_p, _p, _p, _p, _p],
    "mcl_tri_pos_ptr": [_p],
    "mcl_tri_target_ptr": [_p],
    "mcl_tri_move": [_p, _p, _p],
    "mcl_tri_move_vertices": [_p, _p, _p, _i, _i, _p, _p],
    "mcl_hmesh_move_vertex": [_p, _p, _i, _p, _i, _p, _i, _i, _i, _p, _i, _p],
    "mcl_hmesh_delonize": [_p, _p, _i, _p, _i, _p, _i, _i, _p, _i, _p],
    "mcl_hmesh_untangle": [_p, _p, _i, _p, _i, _p, _i, _i, _p],
    "mcl_tri_integrate": [_p, _p, _p, _p],
    "mcl_tri_delonize": [_p, _i, _i, _p, _p],
    "mcl_tri_untangle": [_p, _i, _p, _p],
    "mcl_tri_count_inverted": [_p, _p, _p],
    "mcl_tri_count_non_delaunay": [_p, _p, _p],
    "mcl_tri_step": [_p, _p, _p, _p],
    "mcl_tri_timestep": [_p, _p, _p, _p, _i, _p, _f, _p, _p, _i, _p, _p],
    "mcl_tri_timestep_batch": [_p, _p, _p, _p, _i, _p, _p, _p, _i, _p, _p],
    "mcl_tri_one_ring": [_p, _i, _i, _p, ctypes.POINTER(_i), _p],
    "mcl_tri_one_ring_csr": [_p, _p, _p, _i, _p, _p],
    "mcl_tri_closest": [_p, _p, _p, _p, _p],
    "mcl_grid_create": [_i, _f, _f, _f, _f, _i, _i, _i, _pp],
    "mcl_grid_destroy": [_p],
    "mcl_grid_update": [_p, _p, _p, _p],
    "mcl_grid_index_ptrs": [_p, _pp, _pp, _pp, _pp],
    "mcl_grid_frnn_csr": [_p, _p, _f, _p, _p, _i, _p, _p],
    "mcl_grid_mean_occupancy": [_p, _p, _p],
    "mcl_grid_build": [_p, _p, _p],
    "mcl_grid_sorted_ptr": [_p],
    "mcl_grid_frnn_csr_noalloc": [_p, _i, _f, _f, _f, _f, _i, _i, _i, _f, _p, _p, _p, _i, _p, _p],
    "mcl_tri_set_option": [_p, _i, ctypes.c_longlong],
    "mcl_tri_get_option": [_p, _i, ctypes.POINTER(ctypes.c_longlong)],
    "mcl_grid_set_option": [_p, _i, ctypes.c_longlong],
    "mcl_grid_get_option": [_p, _i, ctypes.POINTER(ctypes.c_longlong)],
    "mcl_tuning_path": [],
    "mcl_tri_kernel_timer": [_p, _i],
    "mcl_tri_kernel_timer_read": [_p, _i, _p, _p],
    "mcl_grid_kernel_timer": [_p, _i],
    "mcl_grid_kernel_timer_read": [_p, _i, _p, _p],
    "mcl_eval_predicates": [_i, _i, _p, _i, _p, _p, _p],
    "mcl_tri_closest_ring": [_p, _i, _i, _p, _p, _p],
    "mcl_tri_move_legalize": [_p, _p, _i, _p, _p],
    "mcl_tri_frnn_graph": [_p, _f, _i, _i, _p, _i, _p, ctypes.POINTER(_i), _p],
    "mcl_bgrid_create": [_i, _p, _p, _pp],
    "mcl_bgrid_destroy": [_p],
    "mcl_bgrid_sizes": [_p, _p, _p, _p],
    "mcl_bgrid_build": [_p, _p, _p],
    "mcl_bgrid_sorted_ptr": [_p],
    "mcl_bgrid_index_ptrs": [_p, _pp, _pp, _pp],
    "mcl_bgrid_frnn_csr": [_p, _p, _p, _i, _p, _p],
    "mcl_bgrid_kernel_timer": [_p, _i],
    "mcl_bgrid_kernel_timer_read": [_p, _i, _p, _p],
    "mcl_tri_is_periodic": [_p, _p],
    "mcl_tri_build": [_p, _i, _i, _p, _pp, _p],
    "mcl_batch_build": [_p, _p, _i, _p, _pp, _p],
    "mcl_state_save": [ctypes.c_char_p, _i, _p, _i, _p, _i, _p, _i, _p, _i],
    "mcl_state_info": [ctypes.c_char_p, _p, _p, _p, _p, _p],
    "mcl_state_load": [ctypes.c_char_p, _p, _p, _p, _p],
    "mcl_tri_save": [_p, ctypes.c_char_p],
    "mcl_tri_load": [ctypes.c_char_p, _pp],
    "mcl_version": [],
    "mcl_device_info": [_p, _i, _p],
}
RESTYPES = {"mcl_tuning_path": ctypes.c_char_p, "mcl_tri_pos_ptr": _p, "mcl_tri_target_ptr": _p, "mcl_grid_sorted_ptr": _p, "mcl_bgrid_sorted_ptr": _p,
            "mcl_version": ctypes.c_char_p}


def lib():
    """Load libmcleap_hip.so; raises if it has not been built."""
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmcleap_hip.so is not built (run `python -m mcleap_amd.build`)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = RESTYPES.get(name, _i)
        _L = L
    return _L


def check(rc, what):
    if rc != MCL_OK:
        raise MclError(rc, what)
    return rc
