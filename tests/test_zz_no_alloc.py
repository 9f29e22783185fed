"""The allocation-free radix sort (b40c_radix_sort_u64_noalloc) against the
allocating entry point (b40c_radix_sort_u64), through the C ABI: two calls in
a row, the second smaller (the shared scratch array keeps the first call's
contents), bit-exact keys and the same selector semantics.  Inputs and
outputs are allocated by the test; the entry point under test allocates
nothing."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(lib, nbytes):
    p = ctypes.c_void_p()
    assert lib.b40c_device_alloc(ctypes.byref(p), int(max(nbytes, 8))) == 0
    return p


def _sorted_by(lib, entry, keys, begin_bit, end_bit):
    n = keys.size
    d0, d1 = _dev(lib, keys.nbytes), _dev(lib, keys.nbytes)
    try:
        assert lib.b40c_memcpy_h2d(d0, ctypes.c_void_p(keys.ctypes.data), keys.nbytes) == 0
        sel = ctypes.c_int32(-1)
        st = getattr(lib, entry)(d0, d1, n, begin_bit, end_bit, ctypes.byref(sel))
        out = np.empty(n, np.uint64)
        if st == 0:
            assert sel.value in (0, 1)
            src = d1 if sel.value else d0
            assert lib.b40c_memcpy_d2h(ctypes.c_void_p(out.ctypes.data), src, out.nbytes) == 0
        return st, out
    finally:
        lib.b40c_device_free(d0)
        lib.b40c_device_free(d1)


def test_noalloc_sort_matches_allocating_sort():
    from back40computing_amd import _lib
    lib = _lib.lib()
    rng = np.random.default_rng(123)
    # Second call smaller than the first; bit ranges with and without
    # trivial digit places.
    for n, begin_bit, end_bit, high in ((3_000_001, 0, 64, 2**64 - 1), (40_000, 0, 44, 2**44 - 1),
                                        (70_001, 0, 64, 2**20)):
        keys = rng.integers(0, high, n, dtype=np.uint64)
        print("noalloc sort n=%d bits=[%d,%d)" % (n, begin_bit, end_bit), flush=True)
        st_ref, ref = _sorted_by(lib, "b40c_radix_sort_u64", keys, begin_bit, end_bit)
        st, got = _sorted_by(lib, "b40c_radix_sort_u64_noalloc", keys, begin_bit, end_bit)
        assert st_ref == 0 and st == 0, _lib.last_error()
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got, np.sort(keys))


def test_noalloc_sort_edges_and_limit():
    from back40computing_amd import _lib
    lib = _lib.lib()
    for n in (0, 1, 2, 2049):
        keys = np.arange(n, 0, -1, dtype=np.uint64)
        st, got = _sorted_by(lib, "b40c_radix_sort_u64_noalloc", keys, 0, 64)
        assert st == 0
        np.testing.assert_array_equal(got, np.sort(keys))
    # Beyond the maximum size: refused before any device work.
    d = _dev(lib, 64)
    sel = ctypes.c_int32()
    try:
        st = lib.b40c_radix_sort_u64_noalloc(d, d, 1 << 32, 0, 64, ctypes.byref(sel))
        assert st == _lib.ERR_UNSUPPORTED
        assert "exceeds" in _lib.last_error()
    finally:
        lib.b40c_device_free(d)
