"""CPU-side checks of the C-ABI library: it builds, loads, exports every
symbol the headers declare, and rejects bad host arguments before touching
the GPU.  No compute calls are made here."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(b40c_[a-z0-9_]+)\s*\(", text))
    return syms


@pytest.fixture(scope="module")
def L():
    from back40computing_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from back40computing_amd import build
        build.build()
    return _lib.lib()


def test_exports_every_header_symbol(L):
    syms = header_symbols()
    assert len(syms) > 20
    for s in sorted(syms):
        assert hasattr(L, s), "missing export " + s


def test_ctypes_signatures_cover_header():
    from back40computing_amd import _lib
    assert header_symbols() == set(_lib.SIGNATURES)


def test_host_argument_validation(L):
    from back40computing_amd import _lib
    out = ctypes.c_void_p()
    # COO edge out of range is rejected on the host.
    rows = np.array([0, 5], np.int32)
    cols = np.array([1, 0], np.int32)
    st = L.b40c_graph_from_coo(3, 2, _lib.ptr(rows), _lib.ptr(cols), ctypes.byref(out))
    assert st == _lib.ERR_INVALID_ARG
    assert "out of range" in _lib.last_error()
    # CSR offsets that disagree with the edge count.
    off = np.array([0, 1, 3], np.int32)
    c2 = np.array([1, 0], np.int32)
    st = L.b40c_graph_from_host_csr(2, 2, _lib.ptr(off), _lib.ptr(c2), ctypes.byref(out))
    assert st == _lib.ERR_INVALID_ARG
    # CSR offsets that decrease, and a column id outside [0, n), are refused
    # on the host (the kernels use both as addresses).
    off = np.array([0, 2, 1, 3], np.int32)
    c3 = np.array([1, 2, 0], np.int32)
    st = L.b40c_graph_from_host_csr(3, 3, _lib.ptr(off), _lib.ptr(c3), ctypes.byref(out))
    assert st == _lib.ERR_INVALID_ARG
    assert "decrease" in _lib.last_error()
    off = np.array([0, 1, 2, 3], np.int32)
    for bad in (3, -1, 1 << 30):
        c4 = np.array([1, bad, 0], np.int32)
        st = L.b40c_graph_from_host_csr(3, 3, _lib.ptr(off), _lib.ptr(c4), ctypes.byref(out))
        assert st == _lib.ERR_INVALID_ARG
        assert "out of range" in _lib.last_error()
    # Null selector for the radix sort.
    assert L.b40c_radix_sort_u64(None, None, 10, 0, 64, None) == _lib.ERR_INVALID_ARG
    # Unknown strategy and null handles.
    assert L.b40c_problem_reset(None, 0, 1.0) == _lib.ERR_INVALID_ARG
    assert L.b40c_enact_search(None, None, 0, 0, 0) == _lib.ERR_INVALID_ARG
    assert L.b40c_enactor_frontier_type(None) == -1
    # Microbenchmark search without a graph or a level count.
    n = ctypes.c_int32()
    assert L.b40c_microbench_search(None, 0, 0, None, None, 0, ctypes.byref(n)) == _lib.ERR_INVALID_ARG
    assert L.b40c_microbench_search(None, 0, 0, None, None, 0, None) == _lib.ERR_INVALID_ARG


def test_strategy_frontier_types_match_reference_numbering():
    from back40computing_amd import _lib
    # test/bfs/test_bfs.cu:72-79
    assert (_lib.EXPAND_CONTRACT, _lib.CONTRACT_EXPAND, _lib.TWO_PHASE, _lib.HYBRID) == (0, 1, 2, 3)
    # csr_problem.cuh:54-61
    assert (_lib.VERTEX_FRONTIERS, _lib.EDGE_FRONTIERS, _lib.MIXED_FRONTIERS) == (0, 1, 2)


def test_cpp_driver_built_and_usage():
    """The C++ facade compiles (the driver is built from it) and the driver
    rejects a bad command line without touching the GPU."""
    import os
    import subprocess
    drv = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_bfs")
    assert os.path.exists(drv), "C++ driver not built"
    assert os.path.exists(os.path.join(os.path.dirname(drv), "test_sort")), "C++ sort driver not built"
    r = subprocess.run([drv], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage:" in r.stderr
