"""GPU parity of the primitives next to the CSR build: consecutive removal
(consecutive_removal::Enactor::Trim — first element of every run, as
test/consecutive_removal/test_consecutive_removal.cu:135-159 computes it),
prefix scan (scan::Enactor::Scan) and the deduplicating COO -> CSR build,
against numpy (diff/cumsum/unique) and the serial BFS oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import back40computing_amd as b4  # noqa: E402


@pytest.fixture(scope="module")
def prim():
    from back40computing_amd import primitives
    return primitives


def runs(rng, n, distinct):
    return np.sort(rng.integers(0, max(1, distinct), n).astype(np.uint64))


TRIM_CASES = [(0, 1), (1, 1), (255, 10), (256, 300), (257, 2), (100_000, 5000), (5_000_003, 1_000_000),
              (1 << 20, 1)]


@pytest.mark.parametrize("n,distinct", TRIM_CASES)
def test_trim_keys_and_values(prim, n, distinct):
    rng = np.random.default_rng(n + distinct)
    k = runs(rng, n, distinct)
    if n > 10:
        k[n // 2] = np.uint64(2 ** 63 + 5)  # a high-bit key inside a run boundary
        k = np.sort(k)
    v = np.arange(n, dtype=np.uint32)
    print("trim n=%d distinct=%d" % (n, distinct), flush=True)
    head = np.ones(n, bool)
    if n > 1:
        head[1:] = k[1:] != k[:-1]
    gk, gv = prim.trim(k, v)
    np.testing.assert_array_equal(gk, k[head])
    np.testing.assert_array_equal(gv, v[head])
    gk2, _ = prim.trim(k)
    np.testing.assert_array_equal(gk2, k[head])


def test_trim_unsorted_keeps_every_run(prim):
    """Not a unique: equal keys that are not adjacent both stay."""
    k = np.array([3, 3, 1, 1, 1, 3, 2, 2, 3], np.uint64)
    gk, _ = prim.trim(k)
    assert gk.tolist() == [3, 1, 3, 2, 3]
    alt = np.tile(np.array([7, 9], np.uint64), 70_000)
    gk, _ = prim.trim(alt)
    np.testing.assert_array_equal(gk, alt)


@pytest.mark.parametrize("n", [0, 1, 257, 1 << 20, 3_000_001])
@pytest.mark.parametrize("exclusive", [True, False])
def test_scan(prim, n, exclusive):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    print("scan n=%d exclusive=%d" % (n, exclusive), flush=True)
    got, total = prim.scan(x, exclusive)
    incl = np.cumsum(x, dtype=np.uint64)
    ref = (incl - x.astype(np.uint64)) if exclusive else incl
    np.testing.assert_array_equal(got, (ref & 0xFFFFFFFF).astype(np.uint32))
    assert total == int(x.astype(np.uint64).sum())


def test_scan_offsets_of_degrees(prim):
    """The CSR use: exclusive scan of degrees gives the row offsets."""
    from back40computing_amd import bfs
    g = bfs.Graph.rmat(14, 16, 0.57, 0.19, 0.19, 2)
    off, _ = g.to_host()
    deg = np.diff(off).astype(np.uint32)
    got, total = prim.scan(deg, True)
    np.testing.assert_array_equal(got, off[:-1].astype(np.uint32))
    assert total == off[-1]


def coo_expected(n, rows, cols, drop, sym):
    r, c = rows.astype(np.int64), cols.astype(np.int64)
    if sym:
        r, c = np.concatenate([r, c]), np.concatenate([c, r])
    if drop:
        keep = r != c
        r, c = r[keep], c[keep]
    key = np.unique(r * n + c)
    rr, cc = key // n, key % n
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, rr + 1, 1)
    return np.cumsum(off).astype(np.int32), cc.astype(np.int32)


@pytest.mark.parametrize("drop,sym", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_coo_unique_build(drop, sym):
    from back40computing_amd import bfs
    rng = np.random.default_rng(3 + 2 * drop + sym)
    n, m = 3000, 40_000
    rows = rng.integers(0, n, m).astype(np.int32)
    cols = rng.integers(0, n, m).astype(np.int32)
    rows[:100] = cols[:100]          # self-loops
    rows[100:2000] = rows[2000:3900]  # duplicates
    cols[100:2000] = cols[2000:3900]
    g = bfs.Graph.from_coo_unique(n, rows, cols, drop_self_loops=bool(drop), symmetrize=bool(sym))
    off, cc = g.to_host()
    eoff, ecc = coo_expected(n, rows, cols, drop, sym)
    np.testing.assert_array_equal(off, eoff)
    np.testing.assert_array_equal(cc, ecc)


def test_coo_unique_bfs_matches_oracle(orc):
    from back40computing_amd import bfs
    rng = np.random.default_rng(12)
    n, m = 20_000, 120_000
    rows = rng.integers(0, n, m).astype(np.int32)
    cols = rng.integers(0, n, m).astype(np.int32)
    g = bfs.Graph.from_coo_unique(n, rows, cols, drop_self_loops=True, symmetrize=True)
    off, cc = g.to_host()
    for s in orc.pick_sources(off, 5, 3):
        ref, _, rsd = orc.bfs(off, cc, int(s), with_pred=False)
        for strat in (b4.DIRECTION_OPTIMIZING, b4.EXPAND_CONTRACT):
            print("coo-unique bfs strategy=%d src=%d" % (strat, s), flush=True)
            depth, pred, st = bfs.bfs(g, int(s), strategy=strat)
            np.testing.assert_array_equal(depth, ref)
            assert orc.check_pred(off, cc, ref, pred) == -1
            assert st["search_depth"] == rsd


def test_bad_arguments(prim):
    from back40computing_amd import B40CError, _lib
    L = _lib.lib()
    cnt = __import__("ctypes").c_int64()
    assert L.b40c_consecutive_removal_u64(None, None, None, None, 10, cnt) == _lib.ERR_INVALID_ARG
    from back40computing_amd import bfs
    with pytest.raises(B40CError):
        bfs.Graph.from_coo_unique(4, [0, 5], [1, 0])


def test_unaligned_pointers_use_the_scalar_path(prim):
    """Device pointers not on a 16-byte boundary take the element-wise
    kernels; results must not change."""
    import ctypes
    from back40computing_amd import _lib
    from back40computing_amd.radix_sort import DeviceArray
    L = _lib.lib()
    rng = np.random.default_rng(8)
    n = 300_001
    x = rng.integers(0, 1000, n).astype(np.uint32)
    d = DeviceArray((n + 4) * 4)
    d.upload(np.concatenate([np.zeros(1, np.uint32), x]))
    off = ctypes.c_void_p(d.ptr.value + 4)
    tot = ctypes.c_uint64()
    _lib.check(L.b40c_scan_u32(off, off, n, 1, ctypes.byref(tot)))
    got = d.download(np.uint32, n + 1)[1:]
    np.testing.assert_array_equal(got, (np.cumsum(x, dtype=np.uint64) - x).astype(np.uint32))
    k = np.cumsum(rng.integers(0, 2, n)).astype(np.uint64)
    kin = DeviceArray((n + 1) * 8)
    kout = DeviceArray((n + 1) * 8)
    kin.upload(np.concatenate([np.zeros(1, np.uint64), k]))
    cnt = ctypes.c_int64()
    _lib.check(L.b40c_consecutive_removal_u64(ctypes.c_void_p(kin.ptr.value + 8), None,
                                              ctypes.c_void_p(kout.ptr.value + 8), None, n, ctypes.byref(cnt)))
    head = np.ones(n, bool)
    head[1:] = k[1:] != k[:-1]
    assert cnt.value == int(head.sum())
    np.testing.assert_array_equal(kout.download(np.uint64, n + 1)[1:cnt.value + 1], k[head])
