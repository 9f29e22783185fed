"""ctypes binding of libb40c_mi355x.so (the C ABI declared in include/b40c_mi355x.h).

The library must have been built (``back40computing_amd.build.build()``);
loading fails loudly otherwise — there is no CPU fallback for any entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libb40c_mi355x.so")

# Status codes (b40c_mi355x.h).
OK = 0
ERR_QUEUE_OVERFLOW = 1
ERR_UNSUPPORTED = 2
ERR_OOM = 3
ERR_INVALID_ARG = 4
ERR_HIP = 5
ERR_ADMISSION = 6

STATUS_NAMES = {
    OK: "OK", ERR_QUEUE_OVERFLOW: "QUEUE_OVERFLOW", ERR_UNSUPPORTED: "UNSUPPORTED",
    ERR_OOM: "OOM", ERR_INVALID_ARG: "INVALID_ARG", ERR_HIP: "HIP", ERR_ADMISSION: "ADMISSION",
}

# Strategies (test/bfs/test_bfs.cu:72-79 numbering, plus two of this library).
EXPAND_CONTRACT = 0
CONTRACT_EXPAND = 1
TWO_PHASE = 2
HYBRID = 3
PERSISTENT = 4
DIRECTION_OPTIMIZING = 5

VERTEX_FRONTIERS = 0
EDGE_FRONTIERS = 1
MIXED_FRONTIERS = 2

MODE_DEFAULT = 0
MODE_FUSED = 1
MODE_ITERATIVE = 2

# graph file formats (include/b40c_mi355x.h B40C_FILE_*)
FILE_MARKET = 0
FILE_METIS = 1
FILE_DIMACS = 2

# b40c_graph_from_coo_unique flags
COO_DROP_SELF_LOOPS = 1
COO_SYMMETRIZE = 2


# b40c_microbench_search modes
MB_GATHER_LOOKUP = 0
MB_CULL = 1


class MbLevel(ctypes.Structure):
    """b40c_mb_level: one iteration of the per-stage microbenchmark search."""
    _fields_ = [
        ("iteration", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("lookup_in", ctypes.c_int64),
        ("bitmask_pass", ctypes.c_int64),
        ("label_pass", ctypes.c_int64),
        ("local_pass", ctypes.c_int64),
        ("discovered", ctypes.c_int64),
        ("gathered", ctypes.c_int64),
        ("gather_checksum", ctypes.c_uint64),
        ("offsets_ms", ctypes.c_float),
        ("bench_lookup_ms", ctypes.c_float),
        ("lookup_ms", ctypes.c_float),
        ("bench_gather_ms", ctypes.c_float),
        ("gather_ms", ctypes.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class B40CError(RuntimeError):
    """A non-OK b40c_status; ``.status`` holds the code."""

    def __init__(self, status, message):
        super().__init__("%s (%d): %s" % (STATUS_NAMES.get(status, "?"), status, message))
        self.status = status


class Tuning(ctypes.Structure):
    _fields_ = [
        ("grid_size", ctypes.c_int32),
        ("cta_gather_threshold", ctypes.c_int32),
        ("wave_gather_threshold", ctypes.c_int32),
        ("levels_per_sync", ctypes.c_int32),
        ("saturation_quit", ctypes.c_int32),
        ("do_alpha", ctypes.c_int32),
        ("do_beta", ctypes.c_int32),
        ("persistent_wgs_per_cu", ctypes.c_int32),
        ("grid_gather_threshold", ctypes.c_int32),
        ("bottom_up_grid_size", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_d = ctypes.c_double
_p_i32 = ctypes.POINTER(ctypes.c_int32)
_p_i64 = ctypes.POINTER(ctypes.c_int64)
_p_d = ctypes.POINTER(ctypes.c_double)
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes).  Every exported entry point of the header.
SIGNATURES = {
    "b40c_last_error_string": (ctypes.c_char_p, []),
    "b40c_device_info": (_i32, [_p_i32, ctypes.c_char_p, _i32]),
    "b40c_graph_from_host_csr": (_i32, [_i32, _i64, _vp, _vp, _pp]),
    "b40c_graph_from_coo": (_i32, [_i32, _i64, _vp, _vp, _pp]),
    "b40c_graph_from_coo_unique": (_i32, [_i32, _i64, _vp, _vp, _i32, _pp]),
    "b40c_consecutive_removal_u64": (_i32, [_vp, _vp, _vp, _vp, _i64, _p_i64]),
    "b40c_scan_u32": (_i32, [_vp, _vp, _i64, _i32, ctypes.POINTER(ctypes.c_uint64)]),
    "b40c_graph_build_rmat": (_i32, [_i32, _i32, _d, _d, _d, _u64, _pp]),
    "b40c_graph_build_lattice3d": (_i32, [_i32, _u64, _vp, _pp]),
    "b40c_graph_info": (_i32, [_vp, _p_i32, _p_i64]),
    "b40c_graph_copy_to_host": (_i32, [_vp, _vp, _vp]),
    "b40c_graph_destroy": (None, [_vp]),
    "b40c_edge_list_read": (_i32, [ctypes.c_char_p, _i32, _i32, _i32, _pp]),
    "b40c_edge_list_info": (_i32, [_vp, _p_i32, _p_i64, _p_i32]),
    "b40c_edge_list_copy": (_i32, [_vp, _vp, _vp]),
    "b40c_graph_from_edge_list": (_i32, [_vp, _pp]),
    "b40c_edge_list_destroy": (None, [_vp]),
    "b40c_graph_read": (_i32, [ctypes.c_char_p, _i32, _i32, _i32, _pp]),
    "b40c_graph_write_metis": (_i32, [_vp, ctypes.c_char_p]),
    "b40c_radix_sort_u64": (_i32, [_vp, _vp, _i64, _i32, _i32, _p_i32]),
    "b40c_radix_sort_pairs_u32": (_i32, [_vp, _vp, _vp, _vp, _i64, _i32, _i32, _p_i32]),
    "b40c_radix_sort": (_i32, [_i32, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _p_i32]),
    "b40c_radix_sort_u64_noalloc": (_i32, [_vp, _vp, _i64, _i32, _i32, _p_i32]),
    "b40c_sort_enactor_create": (_i32, [_pp]),
    "b40c_sort_enactor_sort": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _p_i32]),
    "b40c_sort_enactor_set_policy": (_i32, [_vp, _i32, _i32]),
    "b40c_sort_enactor_destroy": (None, [_vp]),
    "b40c_device_alloc": (_i32, [_pp, _i64]),
    "b40c_device_free": (_i32, [_vp]),
    "b40c_memcpy_h2d": (_i32, [_vp, _vp, _i64]),
    "b40c_memcpy_d2h": (_i32, [_vp, _vp, _i64]),
    "b40c_problem_create": (_i32, [_i32, _i64, _vp, _vp, _i32, _pp]),
    "b40c_problem_create_from_graph": (_i32, [_vp, _i32, _pp]),
    "b40c_problem_reset": (_i32, [_vp, _i32, _d]),
    "b40c_problem_extract": (_i32, [_vp, _vp, _vp]),
    "b40c_problem_stream": (_vp, [_vp]),
    "b40c_problem_device_labels": (_i32, [_vp, _pp, _pp]),
    "b40c_problem_destroy": (None, [_vp]),
    "b40c_enactor_create": (_i32, [_i32, _i32, _pp]),
    "b40c_enactor_set_tuning": (_i32, [_vp, ctypes.POINTER(Tuning)]),
    "b40c_enactor_get_tuning": (_i32, [_vp, ctypes.POINTER(Tuning)]),
    "b40c_enactor_frontier_type": (_i32, [_vp]),
    "b40c_enact_search": (_i32, [_vp, _vp, _i32, _i32, _i32]),
    "b40c_get_statistics": (_i32, [_vp, _p_i64, _p_i32, _p_d]),
    "b40c_enactor_destroy": (None, [_vp]),
    "b40c_gbarrier_benchmark": (_i32, [_i32, _i32, _i32, _p_d]),
    "b40c_gbarrier_check": (_i32, [_i32, _i32, _i32, _p_i64]),
    "b40c_microbench_search": (_i32, [_vp, _i32, _i32, _vp, ctypes.POINTER(MbLevel), _i32, _p_i32]),
    "b40c_multigraph_pack": (_i32, [_vp, _i32, _pp]),
    "b40c_multigraph_info": (_i32, [_vp, _p_i32, _vp, _vp]),
    "b40c_multigraph_destroy": (None, [_vp]),
    "b40c_msbfs_create": (_i32, [_vp, _i64, _i32, _pp]),
    "b40c_msbfs_set_tuning": (_i32, [_vp, _i32, _i32, _i32]),
    "b40c_msbfs_set_admission": (_i32, [_vp, _i32]),
    "b40c_msbfs_set_levels_per_step": (_i32, [_vp, _i32]),
    "b40c_msbfs_set_persistent": (_i32, [_vp, _i32]),
    "b40c_msbfs_submit": (_i32, [_vp, _i32, _i32, _p_i32]),
    "b40c_msbfs_step": (_i32, [_vp, _p_i32, _p_i32]),
    "b40c_msbfs_run": (_i32, [_vp]),
    "b40c_msbfs_poll": (_i32, [_vp, _vp, _i32, _p_i32]),
    "b40c_msbfs_result": (_i32, [_vp, _i32, _vp, _vp, _p_i32]),
    "b40c_msbfs_stream": (_vp, [_vp]),
    "b40c_msbfs_destroy": (None, [_vp]),
}

_lib = None


def lib():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libb40c_mi355x.so is not built (%s); run back40computing_amd.build.build() "
                "(or __graft_entry__.build()) first — there is no CPU fallback" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    msg = lib().b40c_last_error_string()
    return msg.decode() if msg else ""


def check(status):
    if status != OK:
        raise B40CError(status, last_error())
    return status


def ptr(arr):
    """Host pointer of a C-contiguous numpy array."""
    assert arr.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return ctypes.c_void_p(arr.ctypes.data)
