"""Shared helpers for the test suite (graph construction from edge lists)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def csr_from_edges(n, edges, undirected):
    """CSR with ascending columns per row, built with numpy (independent of the oracle)."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    if undirected:
        e = np.concatenate([e, e[:, ::-1]])
    keys = np.unique(e[:, 0] * n + e[:, 1]) if e.size else np.zeros(0, np.int64)
    rows = (keys // n).astype(np.int32)
    cols = (keys % n).astype(np.int32)
    offsets = np.zeros(n + 1, np.int32)
    np.add.at(offsets, rows + 1, 1)
    offsets = np.cumsum(offsets).astype(np.int32)
    return offsets, cols


def hand_graphs():
    with open(os.path.join(GOLDEN, "hand_graphs.json")) as f:
        return json.load(f)["graphs"]


def scipy_depths(offsets, cols, src):
    """Unweighted shortest-path lengths via scipy.sparse.csgraph (-1 = unreachable)."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    n = offsets.size - 1
    a = sp.csr_matrix((np.ones(cols.size, np.int8), cols, offsets), shape=(n, n))
    d = cg.shortest_path(a, method="D", directed=True, unweighted=True, indices=[src])[0]
    out = np.where(np.isinf(d), -1, d).astype(np.int32)
    return out
