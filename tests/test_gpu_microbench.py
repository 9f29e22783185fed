"""GPU parity of the per-stage microbenchmark search (status lookup and
neighbour gather, each run as a bench pass and a real pass per level, as
test/bfs/microbench/enactor_gather_lookup.cuh:262-377 and
enactor_cull.cuh:262-377 alternate them).  Against the serial oracle's
depths: the labels bit for bit, and per iteration the discovered-vertex
count, the gathered-edge count (the degree sum of that depth), the bench
gather's checksum (the sum of those neighbour ids) and the lookup input (the
previous gather).  The bench culls are best-effort, so their survivor counts
are checked against the bounds every cull must keep: no cull drops the last
copy of an undiscovered vertex, and each stage only removes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import back40computing_amd as b4  # noqa: E402,F401


@pytest.fixture(scope="module")
def mb():
    from back40computing_amd import microbench
    return microbench


def expected_levels(off, cols, depth):
    deg = np.diff(off).astype(np.int64)
    rows = np.repeat(np.arange(off.size - 1), deg)
    top = int(depth.max())
    out = []
    i = 0
    prev_gathered = 0
    while True:
        sel = depth == i
        g = int(deg[sel].sum())
        ck = int(cols[sel[rows]].astype(np.uint64).sum(dtype=np.uint64)) if g else 0
        out.append({"iteration": i, "lookup_in": prev_gathered if i else 0, "discovered": int(sel.sum()),
                    "gathered": g, "gather_checksum": ck})
        if g == 0:
            break
        if i + 1 > top:
            out.append({"iteration": i + 1, "lookup_in": g, "discovered": 0, "gathered": 0, "gather_checksum": 0})
            break
        prev_gathered = g
        i += 1
    return out


def check_levels(levels, exp, mode):
    assert len(levels) == len(exp)
    for r, e in zip(levels, exp):
        for k, v in e.items():
            assert r[k] == v, (k, r, e)
        if r["iteration"] == 0:
            assert r["bitmask_pass"] == r["label_pass"] == r["local_pass"] == -1
            continue
        assert r["discovered"] <= r["label_pass"] <= r["bitmask_pass"] <= r["lookup_in"], r
        if mode == "cull":
            assert r["discovered"] <= r["local_pass"] <= r["label_pass"], r
        else:
            assert r["local_pass"] == -1
        assert r["lookup_ms"] > 0 and r["bench_lookup_ms"] > 0
    for r in levels:
        if r["gathered"]:
            assert r["offsets_ms"] > 0 and r["gather_ms"] > 0 and r["bench_gather_ms"] > 0
        else:
            assert r["offsets_ms"] == r["gather_ms"] == r["bench_gather_ms"] == 0


def run_case(orc, mb, g, src, mode):
    off, cols = g.to_host()
    ref, _, rsd = orc.bfs(off, cols, src, with_pred=False)
    print("microbench mode=%s n=%d m=%d src=%d" % (mode, g.nodes, g.edges, src), flush=True)
    depth, levels = mb.search(g, src, mode)
    np.testing.assert_array_equal(depth, ref)
    check_levels(levels, expected_levels(off, cols, ref), mode)
    return levels


@pytest.mark.parametrize("mode", ["gather_lookup", "cull"])
def test_rmat(orc, mb, mode):
    from back40computing_amd import bfs
    g = bfs.Graph.rmat(16, 16, 0.57, 0.19, 0.19, 3)
    off, _ = g.to_host()
    for s in orc.pick_sources(off, 7, 3):
        run_case(orc, mb, g, int(s), mode)
    # A vertex without out-edges: one record, nothing gathered.
    lone = np.flatnonzero(np.diff(off) == 0)
    if lone.size:
        levels = run_case(orc, mb, g, int(lone[0]), mode)
        assert len(levels) == 1 and levels[0]["gathered"] == 0


@pytest.mark.parametrize("mode", ["gather_lookup", "cull"])
def test_lattice(orc, mb, mode):
    from back40computing_amd import bfs
    g = bfs.Graph.lattice3d(24)
    run_case(orc, mb, g, 0, mode)
    run_case(orc, mb, g, g.nodes // 2, mode)


@pytest.mark.parametrize("mode", ["gather_lookup", "cull"])
def test_irregular_directed(orc, mb, mode):
    """Self-loops, duplicate edges, isolated vertices, a directed graph."""
    from back40computing_amd import bfs
    rng = np.random.default_rng(44)
    n, m = 6000, 30000
    r = rng.integers(0, n - 500, m).astype(np.int32)
    c = rng.integers(0, n - 500, m).astype(np.int32)
    c[:800] = r[:800]
    r[800:2000] = r[2000:3200]
    c[800:2000] = c[2000:3200]
    g = bfs.Graph.from_coo(n, r, c)
    for s in (0, 17, n - 1):
        run_case(orc, mb, g, s, mode)


def test_many_levels_and_truncated_records(orc, mb):
    """A long path: the search outlasts a small record buffer; the library
    reports the full iteration count and writes only what fits."""
    from back40computing_amd import _lib, bfs
    n = 3000
    r = np.arange(n - 1, dtype=np.int32)
    g = bfs.Graph.from_coo(n, np.concatenate([r, r + 1]), np.concatenate([r + 1, r]))
    levels = run_case(orc, mb, g, 0, "cull")
    assert len(levels) == n + 1
    recs = (_lib.MbLevel * 10)()
    cnt = ctypes.c_int32()
    _lib.check(_lib.lib().b40c_microbench_search(g.handle, 0, _lib.MB_GATHER_LOOKUP, None, recs, 10,
                                                 ctypes.byref(cnt)))
    assert cnt.value == n + 1
    assert [recs[i].iteration for i in range(10)] == list(range(10))


def test_edgeless_graph(orc, mb):
    """No edges at all: every search is the source alone, one record."""
    from back40computing_amd import bfs
    g = bfs.Graph.from_coo(5, np.zeros(0, np.int32), np.zeros(0, np.int32))
    for mode in ("gather_lookup", "cull"):
        depth, levels = mb.search(g, 3, mode)
        assert depth.tolist() == [-1, -1, -1, 0, -1]
        assert len(levels) == 1 and levels[0]["discovered"] == 1 and levels[0]["gathered"] == 0


def test_bad_arguments(mb):
    from back40computing_amd import B40CError, bfs
    g = bfs.Graph.rmat(8, 4, 0.57, 0.19, 0.19, 1)
    with pytest.raises(B40CError):
        mb.search(g, g.nodes, "cull")
    with pytest.raises(B40CError):
        mb.search(g, -1)
    with pytest.raises(ValueError):
        mb.search(g, 0, "nope")


def test_summary_accounts_every_id(orc, mb):
    from back40computing_amd import bfs
    g = bfs.Graph.rmat(15, 16, 0.57, 0.19, 0.19, 5)
    off, cols = g.to_host()
    ref, _, _ = orc.bfs(off, cols, 0, with_pred=False)
    _, levels = mb.search(g, 0, "cull")
    s = mb.summarize(levels)
    reached = ref >= 0
    assert s["vertices_found"] == int(reached.sum())
    assert s["edges_gathered"] == int(np.diff(off)[reached].sum())
    assert s["ids_looked_up"] == s["edges_gathered"]
    assert 0 < s["local_pass_frac"] <= s["label_pass_frac"] <= s["bitmask_pass_frac"] <= 1
    assert s["lookup"]["gb_s"] > 0 and s["gather"]["gb_s"] > 0
    assert mb.format_table(levels, "cull").count("\n") == len(levels)
