"""CPU oracle for the BFS path — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker (or as
the timed CPU baseline).  The product package ``back40computing_amd`` never
imports it.  The algorithms live in ``oracle.cpp`` (serial C++, each function
citing the reference file:line it restates); this module is a numpy/ctypes
veneer over that library.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle library in place (g++)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_i32, c_i64, c_u64, c_d = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
        L.orc_splitmix64.argtypes = [c_u64]
        L.orc_splitmix64.restype = c_u64
        L.orc_rmat_thresholds.argtypes = [c_d, c_d, c_d, _u32p]
        L.orc_rmat_edges.argtypes = [ctypes.c_int, c_i64, c_i64, _u32p, c_u64, _i32p, _i32p]
        L.orc_build_rmat_csr.argtypes = [ctypes.c_int, ctypes.c_int, c_d, c_d, c_d, c_u64, _i32p, _i32p]
        L.orc_build_rmat_csr.restype = c_i64
        L.orc_lattice_perm.argtypes = [ctypes.c_int, c_u64, _i32p]
        L.orc_build_lattice_csr.argtypes = [ctypes.c_int, c_u64, _i32p, _i32p, _i32p]
        L.orc_build_lattice_csr.restype = c_i64
        L.orc_lattice_depth.argtypes = [ctypes.c_int, _i32p, c_i32, _i32p]
        L.orc_csr_from_coo.argtypes = [c_i32, c_i64, _i32p, _i32p, _i32p, _i32p]
        L.orc_csr_from_coo.restype = c_i64
        L.orc_bfs.argtypes = [c_i32, _i32p, _i32p, c_i32, _i32p, ctypes.c_void_p]
        L.orc_bfs.restype = c_i32
        L.orc_bfs_timed.argtypes = [c_i32, _i32p, _i32p, c_i32, _i32p,
                                    ctypes.POINTER(c_d), ctypes.POINTER(c_d)]
        L.orc_bfs_timed.restype = c_i32
        L.orc_bfs_many.argtypes = [c_i32, _i32p, _i32p, _i32p, c_i32, _i32p, c_i32]
        L.orc_check_pred.argtypes = [c_i32, _i32p, _i32p, _i32p, _i32p]
        L.orc_check_pred.restype = c_i64
        L.orc_edges_visited.argtypes = [c_i32, _i32p, _i32p, ctypes.POINTER(c_i64)]
        L.orc_edges_visited.restype = c_i64
        _i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.orc_histogram.argtypes = [c_i32, _i32p, _i32p, _i32p, c_i32, _i64p, _i64p, _i64p]
        L.orc_histogram.restype = c_i32
        L.orc_pick_sources.argtypes = [c_i32, _i32p, c_u64, c_i32, _i32p]
        L.orc_pick_sources.restype = c_i32
        L.orc_sort_u64.argtypes = [_u64p, c_i64]
        L.orc_stable_sort_pairs_u32.argtypes = [_u32p, _u32p, c_i64]
        _lib = L
    return _lib


def splitmix64(x):
    return int(lib().orc_splitmix64(x))


def rmat_thresholds(a, b, c):
    t = np.zeros(3, np.uint32)
    lib().orc_rmat_thresholds(a, b, c, t)
    return t


def rmat_edges(scale, i0, i1, a=0.57, b=0.19, c=0.19, seed=0):
    t = rmat_thresholds(a, b, c)
    u = np.empty(i1 - i0, np.int32)
    v = np.empty(i1 - i0, np.int32)
    lib().orc_rmat_edges(scale, i0, i1, t, seed, u, v)
    return u, v


def build_rmat_csr(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0):
    n = 1 << scale
    offsets = np.empty(n + 1, np.int32)
    cols = np.empty(2 * edge_factor * n, np.int32)
    m = lib().orc_build_rmat_csr(scale, edge_factor, a, b, c, seed, offsets, cols)
    return offsets, cols[:m].copy()


def lattice_perm(w, seed=0):
    perm = np.empty(w * w * w, np.int32)
    lib().orc_lattice_perm(w, seed, perm)
    return perm


def build_lattice_csr(w, seed=0):
    n = w * w * w
    offsets = np.empty(n + 1, np.int32)
    cols = np.empty(6 * n, np.int32)
    perm = np.empty(n, np.int32)
    m = lib().orc_build_lattice_csr(w, seed, offsets, cols, perm)
    return offsets, cols[:m].copy(), perm


def lattice_depth(w, perm, src):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size, dtype=np.int32)
    d = np.empty(perm.size, np.int32)
    lib().orc_lattice_depth(w, inv, src, d)
    return d


def csr_from_coo(n, rows, cols):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    offsets = np.empty(n + 1, np.int32)
    out = np.empty(rows.size, np.int32)
    lib().orc_csr_from_coo(n, rows.size, rows, cols, offsets, out)
    return offsets, out


def bfs(offsets, cols, src, with_pred=True):
    """Serial reference BFS.  Returns (depth, pred or None, search_depth)."""
    n = offsets.size - 1
    depth = np.empty(n, np.int32)
    pred = np.empty(n, np.int32) if with_pred else None
    sd = lib().orc_bfs(n, offsets, cols, src, depth,
                       pred.ctypes.data if pred is not None else None)
    return depth, pred, int(sd)


def bfs_timed(offsets, cols, src):
    n = offsets.size - 1
    depth = np.empty(n, np.int32)
    wall = ctypes.c_double()
    user = ctypes.c_double()
    sd = lib().orc_bfs_timed(n, offsets, cols, src, depth, ctypes.byref(wall), ctypes.byref(user))
    return depth, int(sd), wall.value, user.value


def bfs_many(offsets, cols, srcs, threads=8):
    n = offsets.size - 1
    srcs = np.ascontiguousarray(srcs, np.int32)
    out = np.empty((srcs.size, n), np.int32)
    lib().orc_bfs_many(n, offsets, cols, srcs, srcs.size, out, threads)
    return out


def check_pred(offsets, cols, ref_depth, pred):
    """-1 when every predecessor passes the reference invariants, else the first bad node."""
    n = offsets.size - 1
    return int(lib().orc_check_pred(n, offsets, cols,
                                    np.ascontiguousarray(ref_depth, np.int32),
                                    np.ascontiguousarray(pred, np.int32)))


def check_pred_many(offsets, cols, ref_depths, preds, threads=16):
    """check_pred for several searches of one graph, `threads` at a time
    (ctypes releases the GIL around each call).  Returns the list of results
    in order."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda dp: check_pred(offsets, cols, dp[0], dp[1]), zip(ref_depths, preds)))


def edges_visited(offsets, labels):
    nv = ctypes.c_int64()
    e = lib().orc_edges_visited(offsets.size - 1, offsets, np.ascontiguousarray(labels, np.int32),
                                ctypes.byref(nv))
    return int(e), int(nv.value)


def histogram(offsets, cols, labels, search_depth):
    """(expanded, unique_expanded, discovered) per depth, each of length
    search_depth + 1 (test_bfs.cu:246-337)."""
    ex = np.zeros(search_depth + 1, np.int64)
    ue = np.zeros(search_depth + 1, np.int64)
    dc = np.zeros(search_depth + 1, np.int64)
    rc = lib().orc_histogram(offsets.size - 1, np.ascontiguousarray(offsets, np.int32),
                             np.ascontiguousarray(cols, np.int32), np.ascontiguousarray(labels, np.int32),
                             int(search_depth), ex, ue, dc)
    if rc != 0:
        raise ValueError("label outside [-1, search_depth)")
    return ex, ue, dc


def pick_sources(offsets, seed, count):
    out = np.empty(count, np.int32)
    got = lib().orc_pick_sources(offsets.size - 1, offsets, seed, count, out)
    return out[:got].copy()


def sort_u64(keys):
    k = np.array(keys, dtype=np.uint64, copy=True)
    lib().orc_sort_u64(k, k.size)
    return k


def stable_sort_pairs_u32(keys, vals):
    k = np.array(keys, dtype=np.uint32, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    lib().orc_stable_sort_pairs_u32(k, v, k.size)
    return k, v
