"""GPU parity: radix sort and CSR builders against the oracle (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rs():
    from back40computing_amd import radix_sort
    return radix_sort


@pytest.fixture(scope="module")
def G():
    from back40computing_amd.bfs import Graph
    return Graph


@pytest.mark.parametrize("n", [0, 1, 2, 63, 2048, 2049, 100003, 1 << 20])
def test_sort_u64_random(rs, orc, n):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 2**64 - 1, n, dtype=np.uint64)
    print("sort_u64 n=%d" % n, flush=True)
    np.testing.assert_array_equal(rs.sort_u64(k), orc.sort_u64(k))


def test_sort_u64_bit_ranges_and_trivial_passes(rs, orc):
    rng = np.random.default_rng(7)
    # 20-bit keys with the same high bits: passes over bits 24..63 are trivial.
    k = (rng.integers(0, 1 << 20, 300000).astype(np.uint64) | np.uint64(0xAB00000000000000))
    np.testing.assert_array_equal(rs.sort_u64(k), orc.sort_u64(k))
    # Sorting only bits [0, 12) keeps the order of keys that tie on those bits.
    k2 = rng.integers(0, 1 << 16, 50000).astype(np.uint64)
    out = rs.sort_u64(k2, 0, 12)
    order = np.argsort(k2 & np.uint64(0xFFF), kind="stable")
    np.testing.assert_array_equal(out, k2[order])
    # All keys identical.
    k3 = np.full(10000, 12345, np.uint64)
    np.testing.assert_array_equal(rs.sort_u64(k3), k3)


@pytest.mark.parametrize("n", [1, 5000, 300001])
def test_sort_pairs_u32_is_stable(rs, orc, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1000, n).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = rs.sort_pairs_u32(keys, vals, 0, 10)
    ok, ov = orc.stable_sort_pairs_u32(keys, vals)
    np.testing.assert_array_equal(sk, ok)
    np.testing.assert_array_equal(sv, ov)


def random_keys(dtype, n, rng):
    dt = np.dtype(dtype)
    if dt.kind == "f":  # finite, both signs, many exact ties
        return (rng.integers(-3000, 3000, n) * 0.25).astype(dt)
    if dt.kind == "i":
        info = np.iinfo(dt)
        return rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
    return rng.integers(0, 5000, n).astype(dt)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32, np.float64, np.uint32, np.uint64])
@pytest.mark.parametrize("with_values", [False, True])
def test_sort_key_types_against_numpy(rs, dtype, with_values):
    """Every KeyTraits key type (radix_sort/sort_utils.cuh:163-183), keys
    alone or with values, against numpy's stable argsort of the same keys:
    the signed and float bit conversions must give the types' own order."""
    for n in (1, 2049, 250001):
        rng = np.random.default_rng(n + np.dtype(dtype).itemsize)
        k = random_keys(dtype, n, rng)
        order = np.argsort(k, kind="stable")
        print("sort %s n=%d values=%s" % (np.dtype(dtype).name, n, with_values), flush=True)
        if with_values:
            v = np.arange(n, dtype=np.uint32)
            sk, sv = rs.sort(k, v)
            np.testing.assert_array_equal(sv, order.astype(np.uint32))
        else:
            sk = rs.sort(k)
        np.testing.assert_array_equal(sk.view(np.dtype(dtype)), k[order])


def test_sort_float_special_values(rs):
    """-NaN, -inf, -1, -0, +0, 1, inf, NaN: the order the reference's float
    conversion gives (every bit of a negative key flipped, the sign bit of a
    positive one), worked out by hand from the bit patterns."""
    bits = np.array([0x3F800000, 0x80000000, 0x7FC00000, 0x00000000,
                     0xFF800000, 0x7F800000, 0xBF800000, 0xFFC00000], np.uint32)
    want = np.array([0xFFC00000, 0xFF800000, 0xBF800000, 0x80000000,
                     0x00000000, 0x3F800000, 0x7F800000, 0x7FC00000], np.uint32)
    got = rs.sort(bits.view(np.float32))
    np.testing.assert_array_equal(got.view(np.uint32), want)
    got64 = rs.sort(np.array([1.0, -0.0, 0.0, -np.inf, -2.5], np.float64))
    np.testing.assert_array_equal(got64.view(np.uint64),
                                  np.array([-np.inf, -2.5, -0.0, 0.0, 1.0], np.float64).view(np.uint64))


def test_sort_enactor_reuses_scratch(rs, orc):
    """One enactor sorting a large then a smaller input (its scratch is kept
    and reused), and the stateless entry point on the same keys."""
    en = rs.Enactor()
    rng = np.random.default_rng(11)
    for n in (300000, 70000):
        k = rng.integers(0, 2**64 - 1, n, dtype=np.uint64)
        db = rs.DoubleBuffer(np.uint64, n)
        db.d_keys[0].upload(k)
        assert en.Sort(db, n) == 0
        np.testing.assert_array_equal(db.d_keys[db.selector].download(np.uint64, n), orc.sort_u64(k))
        np.testing.assert_array_equal(rs.sort_u64(k), orc.sort_u64(k))


@pytest.mark.parametrize("items,wgs", [(8, 1), (8, 4), (8, 8), (16, 1), (16, 3), (16, 4), (16, 8)])
def test_sort_policies(rs, orc, items, wgs):
    """Every launch policy (keys per thread 8/16, 1-8 workgroups per CU)
    sorts bit-exactly: 64-bit keys alone, uint32 pairs (stable) and signed
    64-bit keys with values, at sizes around tile and grid boundaries."""
    from back40computing_amd import _lib
    en = rs.Enactor()
    en.set_policy(items, wgs)
    rng = np.random.default_rng(items * 16 + wgs)
    for n in (1, 4095, 4097, 256 * 16 * 256 * wgs + 17, 1 << 21):
        print("sort policy items=%d wgs=%d n=%d" % (items, wgs, n), flush=True)
        k = rng.integers(0, 2**64 - 1, n, dtype=np.uint64)
        db = rs.DoubleBuffer(np.uint64, n)
        db.d_keys[0].upload(k)
        assert en.Sort(db, n) == 0
        np.testing.assert_array_equal(db.d_keys[db.selector].download(np.uint64, n), orc.sort_u64(k))
        db.d_keys[0].free()
        db.d_keys[1].free()
        k32 = rng.integers(0, 1000, n).astype(np.uint32)
        v = np.arange(n, dtype=np.uint32)
        order = np.argsort(k32, kind="stable")
        db = rs.DoubleBuffer(np.uint32, n, with_values=True)
        db.d_keys[0].upload(k32)
        db.d_values[0].upload(v)
        assert en.Sort(db, n) == 0
        np.testing.assert_array_equal(db.d_keys[db.selector].download(np.uint32, n), k32[order])
        np.testing.assert_array_equal(db.d_values[db.selector].download(np.uint32, n), order.astype(np.uint32))
        for a in db.d_keys + db.d_values:
            a.free()
        ki = rng.integers(-2**63, 2**63 - 1, min(n, 300001), dtype=np.int64)
        sk, sv = rs.sort(ki, np.arange(ki.size, dtype=np.uint32))
        order = np.argsort(ki, kind="stable")
        np.testing.assert_array_equal(sk.view(np.int64), ki[order])
    with pytest.raises(_lib.B40CError):
        en.set_policy(12, 4)
    with pytest.raises(_lib.B40CError):
        en.set_policy(8, 9)
    en.close()


def test_sort_autotuner_small(rs):
    """The sort tuner times every policy on a small input, checks each
    output against numpy's stable sort, and returns one of them."""
    from back40computing_amd import autotune
    space = [(8, 4), (16, 2)]
    best, ms, hist = autotune.tune_sort(200003, "u64", False, repeats=1, space=space)
    assert best in space and ms > 0
    assert all(t is not None for _, t in hist)
    best, ms, hist = autotune.tune_sort(100001, "f32", True, repeats=1, space=space)
    assert best in space and all(t is not None for _, t in hist)


@pytest.mark.parametrize("scale,ef,seed", [(6, 4, 1), (10, 16, 0), (14, 16, 5), (16, 16, 0), (22, 16, 0)])
def test_rmat_csr_bit_exact(G, orc, scale, ef, seed):
    print("rmat scale=%d ef=%d seed=%d" % (scale, ef, seed), flush=True)
    g = G.rmat(scale, ef, 0.57, 0.19, 0.19, seed)
    off, cols = g.to_host()
    ref_off, ref_cols = orc.build_rmat_csr(scale, ef, 0.57, 0.19, 0.19, seed)
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(cols, ref_cols)
    g.close()


def test_rmat_other_probabilities(G, orc):
    g = G.rmat(12, 8, 0.45, 0.15, 0.15, 3)
    off, cols = g.to_host()
    ref_off, ref_cols = orc.build_rmat_csr(12, 8, 0.45, 0.15, 0.15, 3)
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(cols, ref_cols)


@pytest.mark.parametrize("w,seed", [(1, 0), (2, 3), (8, 1), (30, 7)])
def test_lattice_csr_bit_exact(G, orc, w, seed):
    g, perm = G.lattice3d(w, seed, want_perm=True)
    off, cols = g.to_host()
    ref_off, ref_cols, ref_perm = orc.build_lattice_csr(w, seed)
    np.testing.assert_array_equal(perm, ref_perm)
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(cols, ref_cols)


@pytest.mark.parametrize("n,m", [(1, 0), (7, 30), (1000, 20000), (50000, 400000)])
def test_from_coo_matches_reference_stable_order(G, orc, n, m):
    rng = np.random.default_rng(m)
    rows = rng.integers(0, n, m).astype(np.int32)
    cols = rng.integers(0, n, m).astype(np.int32)
    g = G.from_coo(n, rows, cols)
    off, out = g.to_host()
    ref_off, ref_out = orc.csr_from_coo(n, rows, cols)
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(out, ref_out)


def test_host_csr_round_trip(G, orc):
    off, cols = orc.build_rmat_csr(9, 8, seed=2)
    g = G.from_host_csr(off, cols)
    o2, c2 = g.to_host()
    np.testing.assert_array_equal(o2, off)
    np.testing.assert_array_equal(c2, cols)
