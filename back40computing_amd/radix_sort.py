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


KEY_TYPES = {
    np.dtype(np.uint32): 0, np.dtype(np.int32): 1, np.dtype(np.float32): 2,
    np.dtype(np.uint64): 3, np.dtype(np.int64): 4, np.dtype(np.float64): 5,
}


class Enactor:
    """radix_sort::Enactor::Sort(DoubleBuffer&, num_elements) for any key type
    of the reference's KeyTraits (radix_sort/sort_utils.cuh:87-183), keys
    only or with uint32 values.  The enactor owns its scratch (the spine)
    across calls, as the reference Enactor owns its spine storage."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().b40c_sort_enactor_create(ctypes.byref(h)))
        self._h = h

    def set_policy(self, items_per_thread=0, wgs_per_cu=0):
        """Launch policy (0 = the tuned default): keys per thread of a
        downsweep tile (8 or 16) and workgroups per CU (1-8)."""
        check(lib().b40c_sort_enactor_set_policy(self._h, int(items_per_thread), int(wgs_per_cu)))

    def Sort(self, storage, num_elements, begin_bit=0, end_bit=None):
        sel = ctypes.c_int32()
        kt = KEY_TYPES.get(storage.key_dtype)
        if kt is None:
            raise TypeError("key dtype %s is not a radix-sortable type" % storage.key_dtype)
        if end_bit is None:
            end_bit = 8 * storage.key_dtype.itemsize
        v = storage.d_values
        s = storage.selector
        st = lib().b40c_sort_enactor_sort(self._h, kt, storage.d_keys[s].ptr, storage.d_keys[s ^ 1].ptr,
                                          v[s].ptr if v else None, v[s ^ 1].ptr if v else None,
                                          int(num_elements), int(begin_bit), int(end_bit), ctypes.byref(sel))
        if st == 0:
            storage.selector ^= sel.value
        return st

    def close(self):
        if self._h:
            lib().b40c_sort_enactor_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sort(keys, values=None, begin_bit=0, end_bit=None):
    """Host convenience for any supported key dtype: upload, stable sort on
    the GPU, download (keys, or (keys, values))."""
    keys = np.ascontiguousarray(keys)
    db = DoubleBuffer(keys.dtype, keys.size, with_values=values is not None)
    db.d_keys[0].upload(keys)
    if values is not None:
        db.d_values[0].upload(np.ascontiguousarray(values, np.uint32))
    check(Enactor().Sort(db, keys.size, begin_bit, end_bit))
    out_k = db.d_keys[db.selector].download(keys.dtype, keys.size)
    if values is None:
        return out_k
    return out_k, db.d_values[db.selector].download(np.uint32, keys.size)


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
