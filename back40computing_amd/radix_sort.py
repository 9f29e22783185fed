"""GPU LSD radix sort (b40c/radix_sort/enactor.cuh:153-266) over the C ABI.

``DoubleBuffer`` mirrors util::DoubleBuffer (util/multiple_buffering.cuh:65-130):
two device key (and optional value) buffers plus a ``selector`` that names the
buffer holding the sorted result after ``Enactor.Sort``.
"""
import ctypes

import numpy as np

from ._lib import check, lib


class DeviceArray:
    """A raw device allocation owned by Python."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().b40c_device_alloc(ctypes.byref(p), int(nbytes)))
        self.ptr = p
        self.nbytes = int(nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().b40c_memcpy_h2d(self.ptr, ctypes.c_void_p(arr.ctypes.data), arr.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        check(lib().b40c_memcpy_d2h(ctypes.c_void_p(out.ctypes.data), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().b40c_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DoubleBuffer:
    def __init__(self, key_dtype, n, with_values=False):
        self.key_dtype = np.dtype(key_dtype)
        self.n = int(n)
        self.d_keys = [DeviceArray(max(1, n) * self.key_dtype.itemsize) for _ in range(2)]
        self.d_values = [DeviceArray(max(1, n) * 4) for _ in range(2)] if with_values else None
        self.selector = 0


class Enactor:
    """radix_sort::Enactor::Sort(DoubleBuffer&, num_elements)."""

    def Sort(self, storage, num_elements, begin_bit=0, end_bit=None):
        sel = ctypes.c_int32()
        kt = storage.key_dtype
        if end_bit is None:
            end_bit = 8 * kt.itemsize
        if kt == np.uint64 and storage.d_values is None:
            st = lib().b40c_radix_sort_u64(storage.d_keys[0].ptr, storage.d_keys[1].ptr, num_elements,
                                           begin_bit, end_bit, ctypes.byref(sel))
        elif kt == np.uint32 and storage.d_values is not None:
            st = lib().b40c_radix_sort_pairs_u32(storage.d_keys[0].ptr, storage.d_keys[1].ptr,
                                                 storage.d_values[0].ptr, storage.d_values[1].ptr,
                                                 num_elements, begin_bit, end_bit, ctypes.byref(sel))
        else:
            raise TypeError("supported: uint64 keys, or uint32 keys with uint32 values")
        if st == 0:
            storage.selector = sel.value
        return st


def sort_u64(keys, begin_bit=0, end_bit=64):
    """Host convenience: upload, sort on the GPU, download."""
    keys = np.ascontiguousarray(keys, np.uint64)
    db = DoubleBuffer(np.uint64, keys.size)
    db.d_keys[0].upload(keys)
    check(Enactor().Sort(db, keys.size, begin_bit, end_bit))
    return db.d_keys[db.selector].download(np.uint64, keys.size)


def sort_pairs_u32(keys, values, begin_bit=0, end_bit=32):
    keys = np.ascontiguousarray(keys, np.uint32)
    values = np.ascontiguousarray(values, np.uint32)
    db = DoubleBuffer(np.uint32, keys.size, with_values=True)
    db.d_keys[0].upload(keys)
    db.d_values[0].upload(values)
    check(Enactor().Sort(db, keys.size, begin_bit, end_bit))
    return (db.d_keys[db.selector].download(np.uint32, keys.size),
            db.d_values[db.selector].download(np.uint32, keys.size))
