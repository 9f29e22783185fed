"""Host-side mirror of b40c's single-GPU BFS interface over the C ABI.

Class and method names follow the reference so callers port directly:

* ``CsrProblem`` — ``FromHostProblem`` / ``Reset`` / ``ExtractResults``
  (b40c/graph/bfs/csr_problem.cuh:255-750).
* ``EnactorExpandContract``, ``EnactorContractExpand``, ``EnactorTwoPhase``,
  ``EnactorHybrid`` — ``EnactSearch`` / ``EnactFusedSearch`` /
  ``EnactIterativeSearch`` / ``GetStatistics`` / ``GetFrontierType``
  (b40c/graph/bfs/enactor_*.cuh), plus ``EnactorPersistent`` and
  ``EnactorDirectionOptimizing`` which have no reference counterpart.

As in the reference, the CamelCase methods return a status code (0 on
success) instead of raising; the message is available from
``last_error()``.  All work runs on the GPU through libb40c_mi355x.so.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (B40CError, Tuning, check, last_error, lib, ptr,
                   VERTEX_FRONTIERS, EDGE_FRONTIERS, MIXED_FRONTIERS)

__all__ = [
    "Graph", "CsrProblem", "EnactorExpandContract", "EnactorContractExpand", "EnactorTwoPhase",
    "EnactorHybrid", "EnactorPersistent", "EnactorDirectionOptimizing", "VERTEX_FRONTIERS",
    "EDGE_FRONTIERS", "MIXED_FRONTIERS", "device_info", "gbarrier_benchmark",
]


def device_info():
    cus = ctypes.c_int32()
    arch = ctypes.create_string_buffer(64)
    check(lib().b40c_device_info(ctypes.byref(cus), arch, 64))
    return {"num_cus": cus.value, "arch": arch.value.decode()}


def gbarrier_benchmark(grid_size, iterations, light=True):
    """Milliseconds per software grid barrier (test/gbarrier/test_gbarrier.cu).
    light: the persistent kernels' mode (data through agent-scope accesses,
    no cache maintenance); False: agent-scope release/acquire per workgroup."""
    ms = ctypes.c_double()
    check(lib().b40c_gbarrier_benchmark(grid_size, iterations, 1 if light else 0, ctypes.byref(ms)))
    return ms.value


def gbarrier_check(grid_size, iterations, light=True):
    """Number of stale/wrong words seen across the software grid barrier."""
    bad = ctypes.c_int64()
    check(lib().b40c_gbarrier_check(grid_size, iterations, 1 if light else 0, ctypes.byref(bad)))
    return bad.value


class Graph:
    """Device-resident int32 CSR graph (reference counted by the library)."""

    def __init__(self, handle):
        self._h = handle
        n = ctypes.c_int32()
        m = ctypes.c_int64()
        check(lib().b40c_graph_info(self._h, ctypes.byref(n), ctypes.byref(m)))
        self.nodes = n.value
        self.edges = m.value

    @property
    def handle(self):
        return self._h

    @classmethod
    def from_host_csr(cls, row_offsets, column_indices):
        off = np.ascontiguousarray(row_offsets, np.int32)
        cols = np.ascontiguousarray(column_indices, np.int32)
        h = ctypes.c_void_p()
        check(lib().b40c_graph_from_host_csr(off.size - 1, cols.size, ptr(off), ptr(cols), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_coo(cls, nodes, rows, cols):
        """CSR from COO with the reference's stable row-only ordering (CsrGraph::FromCoo)."""
        r = np.ascontiguousarray(rows, np.int32)
        c = np.ascontiguousarray(cols, np.int32)
        h = ctypes.c_void_p()
        check(lib().b40c_graph_from_coo(nodes, r.size, ptr(r), ptr(c), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_coo_unique(cls, nodes, rows, cols, drop_self_loops=False, symmetrize=False):
        """CSR with sorted, duplicate-free rows built on the GPU (sort, consecutive
        removal, offsets); optionally without self-loops and symmetrized."""
        r = np.ascontiguousarray(rows, np.int32)
        c = np.ascontiguousarray(cols, np.int32)
        flags = (_lib.COO_DROP_SELF_LOOPS if drop_self_loops else 0) | (_lib.COO_SYMMETRIZE if symmetrize else 0)
        h = ctypes.c_void_p()
        check(lib().b40c_graph_from_coo_unique(nodes, r.size, ptr(r), ptr(c), flags, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0):
        h = ctypes.c_void_p()
        check(lib().b40c_graph_build_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def lattice3d(cls, width, seed=0, want_perm=False):
        h = ctypes.c_void_p()
        perm = np.empty(width ** 3, np.int32) if want_perm else None
        check(lib().b40c_graph_build_lattice3d(width, seed, ptr(perm) if perm is not None else None,
                                               ctypes.byref(h)))
        g = cls(h)
        return (g, perm) if want_perm else g

    def to_host(self):
        off = np.empty(self.nodes + 1, np.int32)
        cols = np.empty(self.edges, np.int32)
        check(lib().b40c_graph_copy_to_host(self._h, ptr(off), ptr(cols)))
        return off, cols

    def close(self):
        if self._h:
            lib().b40c_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrProblem:
    """CSR problem storage: device graph, depth/predecessor labels, visited
    mask and frontier queues (csr_problem.cuh:68-173)."""

    def __init__(self):
        self._h = None
        self.nodes = 0
        self.edges = 0

    @property
    def handle(self):
        return self._h

    def FromHostProblem(self, stream_from_host, nodes, edges, h_column_indices, h_row_offsets, num_gpus=1):
        """csr_problem.cuh:328-334.  ``stream_from_host`` is accepted for
        signature parity; 288 GB of HBM holds every graph, so the CSR is always
        copied.  Only one GPU is supported."""
        if num_gpus != 1:
            _lib.lib()  # keep the library-loaded contract
            return _lib.ERR_UNSUPPORTED
        cols = np.ascontiguousarray(h_column_indices, np.int32)
        offs = np.ascontiguousarray(h_row_offsets, np.int32)
        h = ctypes.c_void_p()
        st = lib().b40c_problem_create(nodes, edges, ptr(cols), ptr(offs), 0, ctypes.byref(h))
        if st == 0:
            self._close()
            self._h = h
            self.nodes, self.edges = nodes, edges
        return st

    def FromGraph(self, graph):
        """Bind to a device-resident ``Graph`` without a host round trip."""
        h = ctypes.c_void_p()
        st = lib().b40c_problem_create_from_graph(graph.handle, 0, ctypes.byref(h))
        if st == 0:
            self._close()
            self._h = h
            self.nodes, self.edges = graph.nodes, graph.edges
        return st

    def Reset(self, frontier_type, queue_sizing=1.3):
        return lib().b40c_problem_reset(self._h, frontier_type, float(queue_sizing))

    def ExtractResults(self, h_labels=None, h_preds=None):
        """Copies depth labels into ``h_labels`` and predecessors into
        ``h_preds`` (either may be None)."""
        return lib().b40c_problem_extract(self._h, ptr(h_labels) if h_labels is not None else None,
                                          ptr(h_preds) if h_preds is not None else None)

    # pythonic helpers -------------------------------------------------------
    def results(self):
        depth = np.empty(self.nodes, np.int32)
        pred = np.empty(self.nodes, np.int32)
        check(self.ExtractResults(depth, pred))
        return depth, pred

    def stream(self):
        return lib().b40c_problem_stream(self._h)

    def device_labels(self):
        d = ctypes.c_void_p()
        p = ctypes.c_void_p()
        check(lib().b40c_problem_device_labels(self._h, ctypes.byref(d), ctypes.byref(p)))
        return d.value, p.value

    def _close(self):
        if self._h:
            lib().b40c_problem_destroy(self._h)
            self._h = None

    def close(self):
        self._close()

    def __del__(self):
        try:
            self._close()
        except Exception:
            pass


class _Enactor:
    STRATEGY = None

    def __init__(self, DEBUG=False, instrument=False):
        h = ctypes.c_void_p()
        check(lib().b40c_enactor_create(self.STRATEGY, 1 if instrument else 0, ctypes.byref(h)))
        self._h = h
        self.DEBUG = DEBUG

    @property
    def handle(self):
        return self._h

    def GetFrontierType(self):
        return lib().b40c_enactor_frontier_type(self._h)

    def EnactSearch(self, csr_problem, src, max_grid_size=0):
        return lib().b40c_enact_search(self._h, csr_problem.handle, int(src), max_grid_size, _lib.MODE_DEFAULT)

    def EnactFusedSearch(self, csr_problem, src, max_grid_size=0):
        return lib().b40c_enact_search(self._h, csr_problem.handle, int(src), max_grid_size, _lib.MODE_FUSED)

    def EnactIterativeSearch(self, csr_problem, src, max_grid_size=0):
        return lib().b40c_enact_search(self._h, csr_problem.handle, int(src), max_grid_size, _lib.MODE_ITERATIVE)

    def GetStatistics(self):
        """(total_queued, search_depth, avg_duty), enactor_contract_expand.cuh:236-250."""
        q = ctypes.c_int64()
        d = ctypes.c_int32()
        duty = ctypes.c_double()
        check(lib().b40c_get_statistics(self._h, ctypes.byref(q), ctypes.byref(d), ctypes.byref(duty)))
        return q.value, d.value, duty.value

    @property
    def tuning(self):
        t = Tuning()
        check(lib().b40c_enactor_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib().b40c_enactor_set_tuning(self._h, ctypes.byref(t)))

    def close(self):
        if self._h:
            lib().b40c_enactor_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnactorExpandContract(_Enactor):
    """Vertex frontiers; gather + inline cull + compact (enactor_expand_contract.cuh)."""
    STRATEGY = _lib.EXPAND_CONTRACT


class EnactorContractExpand(_Enactor):
    """Edge frontiers; cull at contraction, then expand (enactor_contract_expand.cuh)."""
    STRATEGY = _lib.CONTRACT_EXPAND


class EnactorTwoPhase(_Enactor):
    """Mixed frontiers; contract and expand as separate launches (enactor_two_phase.cuh)."""
    STRATEGY = _lib.TWO_PHASE


class EnactorHybrid(_Enactor):
    """Fused contract-expand for narrow levels, filter/contract/expand for
    wide ones (enactor_hybrid.cuh)."""
    STRATEGY = _lib.HYBRID


class EnactorPersistent(_Enactor):
    """Expand-contract levels in one launch with a software grid barrier."""
    STRATEGY = _lib.PERSISTENT


class EnactorDirectionOptimizing(_Enactor):
    """Top-down expand-contract switching to bottom-up steps on wide levels."""
    STRATEGY = _lib.DIRECTION_OPTIMIZING


ENACTORS = {
    _lib.EXPAND_CONTRACT: EnactorExpandContract,
    _lib.CONTRACT_EXPAND: EnactorContractExpand,
    _lib.TWO_PHASE: EnactorTwoPhase,
    _lib.HYBRID: EnactorHybrid,
    _lib.PERSISTENT: EnactorPersistent,
    _lib.DIRECTION_OPTIMIZING: EnactorDirectionOptimizing,
}


def bfs(graph, src, strategy=_lib.HYBRID, queue_sizing=1.3, mode=_lib.MODE_DEFAULT, tuning=None):
    """One-shot convenience: returns (depth, pred, stats dict).  Raises on error."""
    prob = CsrProblem()
    check(prob.FromGraph(graph))
    en = ENACTORS[strategy]()
    if tuning:
        en.set_tuning(**tuning)
    check(prob.Reset(en.GetFrontierType(), queue_sizing))
    check(lib().b40c_enact_search(en.handle, prob.handle, int(src), 0, mode))
    depth, pred = prob.results()
    q, sd, duty = en.GetStatistics()
    prob.close()
    en.close()
    return depth, pred, {"total_queued": q, "search_depth": sd, "avg_duty": duty}
