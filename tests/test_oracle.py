"""Pins the CPU oracle (test infrastructure) against facts it does not compute
itself: hand-worked golden graphs, scipy's shortest paths, the lattice's
closed-form Manhattan depth, numpy's sorts, and planted mistakes."""
import numpy as np
import pytest

from helpers import csr_from_edges, hand_graphs, scipy_depths


@pytest.mark.parametrize("g", hand_graphs(), ids=lambda g: g["name"])
def test_bfs_hand_golden(orc, g):
    off, cols = csr_from_edges(g["n"], g["edges"], g["undirected"])
    for case in g["cases"]:
        depth, pred, sd = orc.bfs(off, cols, case["src"])
        assert depth.tolist() == case["depth"]
        assert sd == case["search_depth"]
        assert orc.check_pred(off, cols, depth, pred) == -1


def test_bfs_matches_scipy_on_rmat(orc):
    off, cols = orc.build_rmat_csr(12, 8, seed=3)
    for src in orc.pick_sources(off, 7, 6):
        depth, pred, sd = orc.bfs(off, cols, int(src))
        np.testing.assert_array_equal(depth, scipy_depths(off, cols, int(src)))
        assert sd == depth.max() + 1
        assert orc.check_pred(off, cols, depth, pred) == -1


def test_bfs_many_equals_single(orc):
    off, cols = orc.build_rmat_csr(11, 8, seed=5)
    srcs = orc.pick_sources(off, 1, 5)
    many = orc.bfs_many(off, cols, srcs, threads=3)
    for i, s in enumerate(srcs):
        np.testing.assert_array_equal(many[i], orc.bfs(off, cols, int(s))[0])


def test_bfs_timed_equals_untimed(orc):
    off, cols = orc.build_rmat_csr(10, 16, seed=1)
    s = int(orc.pick_sources(off, 2, 1)[0])
    d, sd, wall, user = orc.bfs_timed(off, cols, s)
    d2, _, sd2 = orc.bfs(off, cols, s, with_pred=False)
    np.testing.assert_array_equal(d, d2)
    assert sd == sd2 and wall >= 0 and user >= 0


def test_check_pred_rejects_planted_mistakes(orc):
    off, cols = orc.build_rmat_csr(10, 16, seed=2)
    s = int(orc.pick_sources(off, 3, 1)[0])
    depth, pred, _ = orc.bfs(off, cols, s)
    assert orc.check_pred(off, cols, depth, pred) == -1
    far = np.where(depth >= 2)[0]
    v = int(far[0])
    # parent at the wrong depth (the source itself)
    bad = pred.copy(); bad[v] = s
    assert orc.check_pred(off, cols, depth, bad) == v
    # parent at the right depth but not adjacent
    same_level = np.where(depth == depth[pred[v]])[0]
    nbrs = set(cols[off[v]:off[v + 1]].tolist())
    cand = [int(u) for u in same_level if int(u) not in nbrs]
    if cand:
        bad = pred.copy(); bad[v] = cand[0]
        assert orc.check_pred(off, cols, depth, bad) == v
    # source must carry -2, reached vertices must have a parent
    bad = pred.copy(); bad[s] = -1
    assert orc.check_pred(off, cols, depth, bad) == s
    bad = pred.copy(); bad[v] = -1
    assert orc.check_pred(off, cols, depth, bad) == v


def test_rmat_thresholds_and_stream(orc):
    t = orc.rmat_thresholds(0.57, 0.19, 0.19)
    assert t.tolist() == [int(0.57 * 2**32), int((0.57 + 0.19) * 2**32), int((0.57 + 0.19 + 0.19) * 2**32)]
    # Top-level quadrant frequencies follow (a, b, c, d).
    scale = 8
    u, v = orc.rmat_edges(scale, 0, 200000)
    half = 1 << (scale - 1)
    top = (u >= half).astype(int) * 2 + (v >= half).astype(int)  # 0:a 1:b 2:c 3:d
    freq = np.bincount(top, minlength=4) / top.size
    np.testing.assert_allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.005)
    assert u.min() >= 0 and u.max() < 2**scale and v.max() < 2**scale
    # Counter-based: any sub-range reproduces the same edges.
    u2, v2 = orc.rmat_edges(scale, 1000, 1010)
    np.testing.assert_array_equal(u2, u[1000:1010])
    np.testing.assert_array_equal(v2, v[1000:1010])


def test_rmat_csr_matches_numpy_build(orc):
    scale, ef, seed = 10, 16, 9
    off, cols = orc.build_rmat_csr(scale, ef, seed=seed)
    u, v = orc.rmat_edges(scale, 0, ef << scale, seed=seed)
    keep = u != v
    e = np.stack([u[keep], v[keep]], 1)
    ref_off, ref_cols = csr_from_edges(1 << scale, e, undirected=True)
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(cols, ref_cols)
    # invariants: symmetric, no self loops, strictly increasing columns per row
    rows = np.repeat(np.arange(off.size - 1), np.diff(off))
    assert not np.any(rows == cols)
    fwd = set(zip(rows.tolist(), cols.tolist()))
    assert all((c, r) in fwd for r, c in list(fwd)[:5000])
    for r in range(0, off.size - 1, 97):
        seg = cols[off[r]:off[r + 1]]
        assert np.all(np.diff(seg) > 0)


def test_lattice_closed_form(orc):
    w = 12
    off, cols, perm = orc.build_lattice_csr(w, seed=4)
    n = w ** 3
    assert off[-1] == 6 * w * w * (w - 1)
    assert np.array_equal(np.sort(perm), np.arange(n))
    for src in (0, 5, n // 2, n - 1):
        depth, pred, sd = orc.bfs(off, cols, src)
        np.testing.assert_array_equal(depth, orc.lattice_depth(w, perm, src))
        assert orc.check_pred(off, cols, depth, pred) == -1


def test_lattice_perm_is_rank_of_keys(orc):
    w, seed = 5, 11
    perm = orc.lattice_perm(w, seed)
    n = w ** 3
    bits = int(np.ceil(np.log2(n)))
    low = (1 << bits) - 1
    sm = orc.splitmix64(seed)
    keys = np.array([(orc.splitmix64((sm + o) & (2**64 - 1)) & ~low & (2**64 - 1)) | o for o in range(n)],
                    dtype=np.uint64)
    np.testing.assert_array_equal(perm, np.argsort(np.argsort(keys, kind="stable"), kind="stable"))


def test_csr_from_coo_is_stable_by_row(orc):
    rng = np.random.default_rng(0)
    n, m = 50, 800
    rows = rng.integers(0, n, m).astype(np.int32)
    cols = rng.integers(0, n, m).astype(np.int32)
    off, out = orc.csr_from_coo(n, rows, cols)
    order = np.argsort(rows, kind="stable")
    np.testing.assert_array_equal(out, cols[order])
    np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]))


def test_sort_oracles_match_numpy(orc):
    rng = np.random.default_rng(1)
    k = rng.integers(0, 2**63, 10000, dtype=np.uint64)
    np.testing.assert_array_equal(orc.sort_u64(k), np.sort(k))
    k32 = rng.integers(0, 16, 5000).astype(np.uint32)
    v32 = np.arange(5000, dtype=np.uint32)
    sk, sv = orc.stable_sort_pairs_u32(k32, v32)
    order = np.argsort(k32, kind="stable")
    np.testing.assert_array_equal(sk, k32[order])
    np.testing.assert_array_equal(sv, v32[order])


def test_edges_visited_and_sources(orc):
    off, cols = orc.build_rmat_csr(10, 16, seed=0)
    src = orc.pick_sources(off, 42, 16)
    assert len(set(src.tolist())) == 16
    assert np.all(np.diff(off)[src] > 0)
    np.testing.assert_array_equal(src, orc.pick_sources(off, 42, 16))
    depth, _, _ = orc.bfs(off, cols, int(src[0]))
    e, nv = orc.edges_visited(off, depth)
    mask = depth >= 0
    assert nv == int(mask.sum())
    assert e == int(np.diff(off)[mask].sum())


def test_work_histogram_hand_graph(orc):
    """Histogram (test_bfs.cu:246-337) on a path 0-1-2-3 with leaves 4, 5 on
    vertex 1, from source 0, worked by hand: depths 0 | 1 | 2 4 5 | 3."""
    adj = {0: [1], 1: [0, 2, 4, 5], 2: [1, 3], 3: [2], 4: [1], 5: [1]}
    off = np.cumsum([0] + [len(adj[v]) for v in range(6)]).astype(np.int32)
    cols = np.array([x for v in range(6) for x in adj[v]], np.int32)
    d, _, sd = orc.bfs(off, cols, 0, with_pred=False)
    assert sd == 4
    ex, ue, dc = orc.histogram(off, cols, d, sd)
    assert dc.tolist() == [1, 1, 3, 1, 0]
    assert ex.tolist() == [1, 1, 4, 4, 1]   # degree sums of the previous depth
    assert ue.tolist() == [1, 1, 4, 2, 1]   # {1}, {0,2,4,5}, {1,3}, {2}


def test_work_histogram_invariants(orc):
    """On an RMAT graph: discovered sums to the vertices reached, expanded
    (after level 0) to the edges visited, and every vertex discovered at
    depth d+1 is among the unique endpoints expanded from depth d."""
    off, cols = orc.build_rmat_csr(12, 8, seed=3)
    src = int(orc.pick_sources(off, 4, 1)[0])
    d, _, sd = orc.bfs(off, cols, src, with_pred=False)
    ex, ue, dc = orc.histogram(off, cols, d, sd)
    nv = int((d >= 0).sum())
    ev, _ = orc.edges_visited(off, d)
    assert dc.sum() == nv
    assert ex[1:].sum() == ev
    assert np.all(ue[1:] <= ex[1:])
    assert np.all(dc[1:] <= ue[1:])


def test_bench_source_picker_matches_oracle(orc):
    """The bench and the batched/served workloads pick sources with
    back40computing_amd.sources.pick_sources; the GPU tests with the
    oracle's orc_pick_sources.  Both must return the same sources (same
    seeds, same order), including graphs with many isolated vertices and
    requests for more sources than there are candidates."""
    import numpy as np
    from back40computing_amd.sources import pick_sources
    off16, _ = orc.build_rmat_csr(16, 16, 0.57, 0.19, 0.19, 0)
    off10, _ = orc.build_rmat_csr(10, 4, 0.57, 0.19, 0.19, 9)
    sparse = np.array([0, 0, 2, 2, 2, 3, 3, 3, 4], np.int32)  # 3 vertices with edges out of 8
    for off, seed, count in ((off16, 1, 64), (off16, 11, 300), (off10, 3, 50), (sparse, 5, 3), (sparse, 5, 10)):
        a = pick_sources(off, seed, count)
        b = orc.pick_sources(off, seed, count)
        np.testing.assert_array_equal(a, b)
        assert np.all(np.diff(off)[a] > 0) and len(set(a.tolist())) == a.size
    assert pick_sources(sparse, 5, 10).size == 3


def test_autotuner_label_verifier(orc):
    """The autotuner's own BFS-label check (product code, not the oracle)
    accepts the oracle's labels and rejects each planted mistake: a depth
    off by one, a predecessor that is not a neighbour, a predecessor on the
    wrong level, a reachable vertex left unreached, an unreachable vertex
    labelled, and a wrong source label."""
    import numpy as np
    from back40computing_amd.autotune import sorted_edge_keys, verify_bfs_labels
    graphs = [orc.build_rmat_csr(10, 8, 0.57, 0.19, 0.19, 2), orc.build_lattice_csr(6, 1)[:2]]
    for off, cols in graphs:
        keys = sorted_edge_keys(off, cols)
        src = int(orc.pick_sources(off, 4, 1)[0])
        depth, pred, _ = orc.bfs(off, cols, src, with_pred=True)
        assert verify_bfs_labels(off, cols, src, depth, pred, keys)
        assert verify_bfs_labels(off, cols, src, depth, pred)  # builds its own keys
        v = int(np.flatnonzero(depth == depth.max())[0])
        bad = depth.copy(); bad[v] += 1
        assert not verify_bfs_labels(off, cols, src, bad, pred, keys)
        bad = depth.copy(); bad[v] -= 1
        assert not verify_bfs_labels(off, cols, src, bad, pred, keys)
        nbrs = set(cols[off[v]:off[v + 1]].tolist())
        stranger = next(u for u in range(off.size - 1) if u not in nbrs and u != v and depth[u] == depth[v] - 1) \
            if any(depth[u] == depth[v] - 1 and u not in nbrs for u in range(off.size - 1)) else None
        if stranger is not None:
            bp = pred.copy(); bp[v] = stranger
            assert not verify_bfs_labels(off, cols, src, depth, bp, keys)
            # The predecessor-only mode (depth already accepted) still
            # catches a predecessor that is not a neighbour.
            assert not verify_bfs_labels(off, cols, src, depth, bp, keys, depth_verified=True)
        same_level = [u for u in nbrs if depth[u] == depth[v]]
        if same_level:
            bp = pred.copy(); bp[v] = same_level[0]
            assert not verify_bfs_labels(off, cols, src, depth, bp, keys)
            assert not verify_bfs_labels(off, cols, src, depth, bp, keys, depth_verified=True)
        assert verify_bfs_labels(off, cols, src, depth, pred, keys, depth_verified=True)
        bp = pred.copy(); bp[v] = -1
        assert not verify_bfs_labels(off, cols, src, depth, bp, keys, depth_verified=True)
        bad, bp = depth.copy(), pred.copy()
        bad[v], bp[v] = -1, -1
        assert not verify_bfs_labels(off, cols, src, bad, bp, keys)
        bp = pred.copy(); bp[src] = -1
        assert not verify_bfs_labels(off, cols, src, depth, bp, keys)
    # An isolated vertex labelled as reached.
    off = np.array([0, 1, 2, 2], np.int32)
    cols = np.array([1, 0], np.int32)
    depth = np.array([0, 1, -1], np.int32)
    pred = np.array([-2, 0, -1], np.int32)
    assert verify_bfs_labels(off, cols, 0, depth, pred)
    assert not verify_bfs_labels(off, cols, 0, np.array([0, 1, 2], np.int32), np.array([-2, 0, 1], np.int32))
