"""Seeded source selection for benchmark workloads.

Candidate k is splitmix64(splitmix64(seed ^ 0x5EED) + k) mod n; a candidate
is accepted when its out-degree is positive and it was not chosen before.
(The reference re-rolls sources that reach fewer than 5 vertices,
test/bfs/test_bfs.cu:444-445, 662-666; the workloads here ask for nonzero
degree instead.)
"""
import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def pick_sources(row_offsets, seed, count):
    off = np.asarray(row_offsets)
    n = off.size - 1
    deg = np.diff(off)
    smix = splitmix64(seed ^ 0x5EED)
    out = []
    used = set()
    k = 0
    limit = 64 * n + 1024
    while len(out) < count and k < limit:
        c = splitmix64((smix + k) & _M64) % n
        if deg[c] > 0 and c not in used:
            used.add(c)
            out.append(c)
        k += 1
    return np.array(out, np.int32)
