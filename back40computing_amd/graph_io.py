"""Graph files: MatrixMarket, METIS and DIMACS readers and the METIS writer.

The reference's test driver reads these as graph types ``market``, ``metis``
and ``dimacs`` (test/bfs/test_bfs.cu:976-1037) through
b40c/graph/builder/{market,metis,dimacs}.cuh.  Parsing runs in
libb40c_mi355x.so on the host (``b40c_edge_list_*``, no GPU needed); the
device CSR is then built by the GPU stable sort by row (COO files) or
uploaded as read (METIS).
"""
import ctypes

import numpy as np

from ._lib import FILE_DIMACS, FILE_MARKET, FILE_METIS, check, lib, ptr

FORMATS = {"market": FILE_MARKET, "metis": FILE_METIS, "dimacs": FILE_DIMACS}


def _fmt(fmt):
    return FORMATS[fmt] if isinstance(fmt, str) else int(fmt)


def _path(path):
    return None if path is None else str(path).encode()


class EdgeList:
    """A parsed graph file: COO edges in the reference readers' order, or a
    CSR for METIS (ReadMarketStream / ReadMetisStream / ReadDimacsStream)."""

    def __init__(self, path, fmt, undirected=False, splice=0):
        h = ctypes.c_void_p()
        check(lib().b40c_edge_list_read(_path(path), _fmt(fmt), int(bool(undirected)), int(splice),
                                        ctypes.byref(h)))
        self._h = h
        n, m, csr = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        check(lib().b40c_edge_list_info(h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(csr)))
        self.nodes, self.edges, self.is_csr = n.value, m.value, bool(csr.value)

    def arrays(self):
        """(rows, cols) for a COO list, (row_offsets, cols) for METIS."""
        a = np.empty(self.nodes + 1 if self.is_csr else self.edges, np.int32)
        b = np.empty(self.edges, np.int32)
        check(lib().b40c_edge_list_copy(self._h, ptr(a), ptr(b)))
        return a, b

    def to_graph(self):
        from .bfs import Graph
        h = ctypes.c_void_p()
        check(lib().b40c_graph_from_edge_list(self._h, ctypes.byref(h)))
        return Graph(h)

    def close(self):
        if self._h:
            lib().b40c_edge_list_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_graph(path, fmt, undirected=False, splice=0):
    """Device graph from a file (BuildMarketGraph / BuildMetisGraph /
    BuildDimacsGraph, then the device CSR)."""
    from .bfs import Graph
    h = ctypes.c_void_p()
    check(lib().b40c_graph_read(_path(path), _fmt(fmt), int(bool(undirected)), int(splice), ctypes.byref(h)))
    return Graph(h)


def write_metis(graph, path):
    """METIS text of a device graph (WriteMetisStream, metis.cuh:220-245)."""
    check(lib().b40c_graph_write_metis(graph.handle, _path(path)))
