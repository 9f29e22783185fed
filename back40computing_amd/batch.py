"""Batched and served multi-source BFS over several resident graphs.

``MultiGraph`` packs device graphs into one CSR; ``MultiSourceBFS`` runs many
single-source queries over it as one bit-parallel traversal (up to 64 queries
per graph per level step), admitting queries at level boundaries as graph
slots and label memory free up (continuous batching).  All work runs in
libb40c_mi355x.so (``b40c_msbfs_*`` in include/b40c_mi355x.h).
"""
import ctypes

import numpy as np

from ._lib import check, lib, ptr


class MultiGraph:
    def __init__(self, graphs):
        arr = (ctypes.c_void_p * len(graphs))(*[g.handle for g in graphs])
        h = ctypes.c_void_p()
        check(lib().b40c_multigraph_pack(arr, len(graphs), ctypes.byref(h)))
        self._h = h
        self.ngraphs = len(graphs)
        self.vbase = np.empty(self.ngraphs + 1, np.int32)
        self.ebase = np.empty(self.ngraphs + 1, np.int64)
        check(lib().b40c_multigraph_info(h, None, ptr(self.vbase), ptr(self.ebase)))

    @property
    def handle(self):
        return self._h

    def nodes(self, g):
        return int(self.vbase[g + 1] - self.vbase[g])

    def edges(self, g):
        return int(self.ebase[g + 1] - self.ebase[g])

    def close(self):
        if self._h:
            lib().b40c_multigraph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiSourceBFS:
    def __init__(self, multigraph, label_budget_bytes=1 << 32, max_queries=1 << 20):
        h = ctypes.c_void_p()
        check(lib().b40c_msbfs_create(multigraph.handle, int(label_budget_bytes), int(max_queries), ctypes.byref(h)))
        self._h = h
        self.mg = multigraph

    def set_tuning(self, alpha=0, beta=0, grid=0):
        check(lib().b40c_msbfs_set_tuning(self._h, alpha, beta, grid))

    def set_admission(self, cohort_gap):
        """Levels between admissions while queries run (1: every boundary)."""
        check(lib().b40c_msbfs_set_admission(self._h, int(cohort_gap)))

    def set_levels_per_step(self, levels):
        """Levels per step: >= 1 replays a captured hipGraph of that many
        levels with one synchronisation; 0 launches one level per step."""
        check(lib().b40c_msbfs_set_levels_per_step(self._h, int(levels)))

    def set_persistent(self, on):
        """Run each step's levels in one persistent launch (grid barriers
        between the level's steps) instead of launches or a captured graph."""
        check(lib().b40c_msbfs_set_persistent(self._h, 1 if on else 0))

    def submit(self, graph, src):
        q = ctypes.c_int32()
        check(lib().b40c_msbfs_submit(self._h, int(graph), int(src), ctypes.byref(q)))
        return q.value

    def step(self):
        r = ctypes.c_int32()
        w = ctypes.c_int32()
        check(lib().b40c_msbfs_step(self._h, ctypes.byref(r), ctypes.byref(w)))
        return r.value, w.value

    def run(self):
        check(lib().b40c_msbfs_run(self._h))

    def poll(self, max_ids=1 << 12):
        """Ids of finished queries not returned before (at most max_ids per call)."""
        if getattr(self, "_poll_buf", None) is None or self._poll_buf.size < max_ids:
            self._poll_buf = np.empty(max_ids, np.int32)
        c = ctypes.c_int32()
        check(lib().b40c_msbfs_poll(self._h, ptr(self._poll_buf), max_ids, ctypes.byref(c)))
        out = self._poll_buf[:c.value].tolist()
        if c.value == max_ids:
            out += self.poll(max_ids)
        return out

    def result(self, qid, graph, want_pred=True):
        n = self.mg.nodes(graph)
        depth = np.empty(n, np.int32)
        pred = np.empty(n, np.int32) if want_pred else None
        sd = ctypes.c_int32()
        check(lib().b40c_msbfs_result(self._h, int(qid), ptr(depth), ptr(pred) if pred is not None else None,
                                      ctypes.byref(sd)))
        return depth, pred, sd.value

    def release(self, qid):
        """Frees a finished query's label slot without copying its labels."""
        sd = ctypes.c_int32()
        check(lib().b40c_msbfs_result(self._h, int(qid), None, None, ctypes.byref(sd)))
        return sd.value

    def stream(self):
        return lib().b40c_msbfs_stream(self._h)

    def close(self):
        if self._h:
            lib().b40c_msbfs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
