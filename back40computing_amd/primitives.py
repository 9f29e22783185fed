"""Consecutive removal and prefix scan on the GPU (libb40c_mi355x.so).

* ``ConsecutiveRemoval.Trim`` — consecutive_removal::Enactor::Trim
  (b40c/consecutive_removal/enactor.cuh:119-126): keeps the first key (and
  value) of every run of equal consecutive keys.
* ``Scan.Scan`` — scan::Enactor::Scan (b40c/scan/enactor.cuh:122-136):
  exclusive or inclusive prefix sum of unsigned 32-bit values.

``trim`` and ``scan`` are host-array conveniences around them.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .radix_sort import DeviceArray, DoubleBuffer


class ConsecutiveRemoval:
    def Trim(self, storage, num_elements):
        """Reads d_keys[selector] (and d_values[selector]), writes the
        compacted run heads to the other buffer, flips the selector and
        returns the compacted count."""
        assert storage.key_dtype == np.uint64
        src, dst = storage.selector, storage.selector ^ 1
        vi = storage.d_values[src].ptr if storage.d_values else None
        vo = storage.d_values[dst].ptr if storage.d_values else None
        cnt = ctypes.c_int64()
        check(lib().b40c_consecutive_removal_u64(storage.d_keys[src].ptr, vi, storage.d_keys[dst].ptr, vo,
                                                 int(num_elements), ctypes.byref(cnt)))
        storage.selector = dst
        return cnt.value


class Scan:
    def Scan(self, d_dest, d_src, num_elements, exclusive=True):
        """Prefix sum of num_elements uint32 from d_src into d_dest (may be
        the same array); returns the exact total."""
        tot = ctypes.c_uint64()
        check(lib().b40c_scan_u32(d_src.ptr, d_dest.ptr, int(num_elements), 1 if exclusive else 0,
                                  ctypes.byref(tot)))
        return tot.value


def trim(keys, values=None):
    keys = np.ascontiguousarray(keys, np.uint64)
    n = keys.size
    buf = DoubleBuffer(np.uint64, n, with_values=values is not None)
    buf.d_keys[0].upload(keys)
    if values is not None:
        buf.d_values[0].upload(np.ascontiguousarray(values, np.uint32))
    m = ConsecutiveRemoval().Trim(buf, n)
    k = buf.d_keys[buf.selector].download(np.uint64, m)
    v = buf.d_values[buf.selector].download(np.uint32, m) if values is not None else None
    return k, v


def scan(values, exclusive=True):
    values = np.ascontiguousarray(values, np.uint32)
    d = DeviceArray(max(1, values.nbytes))
    d.upload(values)
    total = Scan().Scan(d, d, values.size, exclusive)
    return d.download(np.uint32, values.size), total
