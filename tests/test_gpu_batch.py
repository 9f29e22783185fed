"""GPU parity of the batched / served multi-source engine.

Every query's depth labels must equal the serial oracle's on its own graph,
predecessors must pass the reference's invariants (test/bfs/test_bfs.cu:
387-438), and a query's labels must be identical whether it runs alone or in
a mix with others (continuous batching, slot reuse, memory admission)."""
import numpy as np
import pytest

from helpers import csr_from_edges

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from back40computing_amd import batch, bfs
    return bfs, batch


@pytest.fixture(scope="module")
def pool(mods, orc):
    """Six symmetric graphs: RMAT of several scales and a lattice."""
    bfs, batch = mods
    specs = [("rmat", 8, 1), ("rmat", 11, 2), ("rmat", 13, 3), ("lattice", 12, 4), ("rmat", 10, 5), ("rmat", 6, 6)]
    graphs, host = [], []
    for kind, p, seed in specs:
        g = bfs.Graph.rmat(p, 16, 0.57, 0.19, 0.19, seed) if kind == "rmat" else bfs.Graph.lattice3d(p, seed)
        graphs.append(g)
        host.append(g.to_host())
    mg = batch.MultiGraph(graphs)
    return mg, host


def check_query(orc, host, g, src, depth, pred, sd):
    off, cols = host[g]
    ref, _, rsd = orc.bfs(off, cols, int(src), with_pred=False)
    np.testing.assert_array_equal(depth, ref)
    assert orc.check_pred(off, cols, ref, pred) == -1
    assert sd == rsd


def run_all(eng, mg, queries):
    ids = {}
    for g, s in queries:
        ids[eng.submit(g, s)] = (g, s)
    eng.run()
    done = eng.poll()
    assert sorted(done) == sorted(ids)
    return {q: (ids[q], eng.result(q, ids[q][0])) for q in done}


def sources(orc, host, g, count, seed):
    return [int(x) for x in orc.pick_sources(host[g][0], seed, count)]


# Levels per step: 0 launches each level's kernels, k >= 1 replays a
# captured hipGraph of k levels (queries finishing mid-replay retire on the
# device).
LEVELS_PER_STEP = [pytest.param(0, id="launch"), pytest.param(1, id="graph1"), pytest.param(4, id="graph4"),
                   pytest.param(-1, id="persistent1"), pytest.param(-3, id="persistent3")]


def set_form(eng, lps):
    """lps >= 0: launches (0) or a captured graph of lps levels; lps < 0: the
    persistent form with -lps levels per step."""
    if lps < 0:
        eng.set_persistent(True)
        eng.set_levels_per_step(-lps)
    else:
        eng.set_levels_per_step(lps)


@pytest.mark.parametrize("lps", LEVELS_PER_STEP)
def test_batch_all_graphs(mods, orc, pool, lps):
    bfs, batch = mods
    mg, host = pool
    queries = [(g, s) for g in range(mg.ngraphs) for s in sources(orc, host, g, 4, 10 + g)]
    eng = batch.MultiSourceBFS(mg, label_budget_bytes=1 << 28)
    set_form(eng, lps)
    print("batch all graphs levels_per_step=%d" % lps, flush=True)
    res = run_all(eng, mg, queries)
    for q, ((g, s), (depth, pred, sd)) in res.items():
        check_query(orc, host, g, s, depth, pred, sd)


@pytest.mark.parametrize("lps", [pytest.param(0, id="launch"), pytest.param(3, id="graph3"),
                                 pytest.param(-2, id="persistent2")])
def test_alone_equals_in_mix(mods, orc, pool, lps):
    """The mix runs in the given launch form, each query alone in the other."""
    bfs, batch = mods
    mg, host = pool
    mix = [(g, s) for g in range(mg.ngraphs) for s in sources(orc, host, g, 6, 20 + g)]
    eng = batch.MultiSourceBFS(mg, label_budget_bytes=1 << 28)
    set_form(eng, lps)
    in_mix = {(g, s): r for q, ((g, s), r) in run_all(eng, mg, mix).items()}
    for g, s in mix[::5]:
        alone_eng = batch.MultiSourceBFS(mg, label_budget_bytes=1 << 28)
        alone_eng.set_levels_per_step(0 if lps else 3)
        (_, (depth, pred, sd)), = run_all(alone_eng, mg, [(g, s)]).values()
        md, mp, msd = in_mix[(g, s)]
        np.testing.assert_array_equal(depth, md)
        assert sd == msd
        check_query(orc, host, g, s, depth, pred, sd)


@pytest.mark.parametrize("lps", LEVELS_PER_STEP)
def test_continuous_batching_and_slot_reuse(mods, orc, pool, lps):
    """Queries arrive while others are in flight; more than 64 on one graph
    forces slot reuse; a small label budget forces memory admission."""
    bfs, batch = mods
    mg, host = pool
    g_big = 2
    budget = 8 * (3 * mg.nodes(g_big))  # room for about three queries on the largest graph
    eng = batch.MultiSourceBFS(mg, label_budget_bytes=budget)
    set_form(eng, lps)
    waves = [[(g_big, s) for s in sources(orc, host, g_big, 5, 31)],
             [(1, s) for s in sources(orc, host, 1, 70, 32)],
             [(3, s) for s in sources(orc, host, 3, 6, 33)] + [(5, s) for s in sources(orc, host, 5, 3, 34)]]
    # Ids are reused once a result is collected, so each id maps to its
    # query only until then.
    ids = {}
    finished = []
    submitted = 0

    def collect():
        for q in eng.poll():
            g, s = ids.pop(q)
            finished.append(((g, s), eng.result(q, g)))

    for wave in waves:
        for g, s in wave:
            ids[eng.submit(g, s)] = (g, s)
            submitted += 1
        for _ in range(2):
            eng.step()
        collect()
    while True:
        r, w = eng.step()
        collect()
        if r == 0 and w == 0:
            break
    assert len(finished) == submitted and not ids
    for (g, s), (depth, pred, sd) in finished:
        check_query(orc, host, g, s, depth, pred, sd)


def test_admission_errors(mods, orc, pool):
    bfs, batch = mods
    from back40computing_amd import B40CError, _lib
    mg, host = pool
    eng = batch.MultiSourceBFS(mg, label_budget_bytes=64)
    with pytest.raises(B40CError) as ei:
        eng.submit(2, 0)
    assert ei.value.status == _lib.ERR_ADMISSION
    with pytest.raises(B40CError):
        eng.submit(0, mg.nodes(0))
    # A directed graph is refused by the batched engine.
    off, cols = csr_from_edges(4, [[0, 1], [1, 2], [2, 0], [2, 3]], undirected=False)
    d = bfs.Graph.from_host_csr(off, cols)
    dm = batch.MultiGraph([d])
    with pytest.raises(B40CError) as ei:
        batch.MultiSourceBFS(dm, label_budget_bytes=1 << 20)
    assert ei.value.status == _lib.ERR_UNSUPPORTED


def test_admission_policy_does_not_change_labels(mods, orc, pool):
    """Cohort admission (the default gap) and admission at every level
    boundary give the same labels for queries that arrive over time."""
    bfs, batch = mods
    mg, host = pool
    arrivals = [[(g, s) for s in sources(orc, host, g, 3, 40 + g)] for g in range(mg.ngraphs)]
    results = []
    for gap in (1, 6):
        eng = batch.MultiSourceBFS(mg, label_budget_bytes=1 << 28)
        eng.set_admission(gap)
        # Ids are reused once a result is collected, so each id maps to its
        # query only until then.
        ids, out = {}, {}
        submitted = 0
        for wave in arrivals:  # one graph's queries per level boundary
            for g, s in wave:
                ids[eng.submit(g, s)] = (g, s)
                submitted += 1
            eng.step()
            for q in eng.poll():
                key = ids.pop(q)
                out[key] = eng.result(q, key[0])[0]
        eng.run()
        for q in eng.poll():
            key = ids.pop(q)
            out[key] = eng.result(q, key[0])[0]
        assert len(out) == submitted and not ids
        results.append(out)
    for key, depth in results[0].items():
        np.testing.assert_array_equal(depth, results[1][key])
        g, s = key
        ref, _, _ = orc.bfs(host[g][0], host[g][1], int(s), with_pred=False)
        np.testing.assert_array_equal(depth, ref)
    with pytest.raises(Exception):
        batch.MultiSourceBFS(mg, label_budget_bytes=1 << 20).set_admission(0)


@pytest.mark.parametrize("gap,lps", [(1, 0), (6, 0), (1, 2), (1, -1)],
                         ids=["gap1-launch", "gap6-launch", "gap1-graph2", "gap1-persistent1"])
def test_staggered_sessions_with_idle_gaps(mods, orc, gap, lps):
    """Long continuous run: one query arrives at every level boundary, the
    engine drains to idle between sessions and restarts, slots are reused.
    Per-level state that is cleared from the frontier queues (mask arrays,
    frontier bitmaps, the pull's graph table) must never leak from one
    query, level or session into another."""
    bfs, batch = mods
    specs = [("rmat", 15, 61), ("rmat", 9, 62), ("lattice", 10, 63), ("rmat", 12, 64)]
    graphs, host = [], []
    for kind, p, seed in specs:
        g = bfs.Graph.rmat(p, 16, 0.57, 0.19, 0.19, seed) if kind == "rmat" else bfs.Graph.lattice3d(p, seed)
        graphs.append(g)
        host.append(g.to_host())
    mg = batch.MultiGraph(graphs)
    eng = batch.MultiSourceBFS(mg, label_budget_bytes=1 << 28)
    eng.set_admission(gap)
    set_form(eng, lps)
    rng = np.random.default_rng(70 + gap)
    ids, out = {}, []
    submitted = 0

    def collect():
        for q in eng.poll():
            g, s = ids.pop(q)  # the id may be reused by the next submission
            out.append(((g, s), eng.result(q, g)))

    for session in range(3):
        arrivals = [int(rng.integers(0, len(specs))) for _ in range(40)]
        for k, g in enumerate(arrivals):
            s = sources(orc, host, g, 1, 1000 * session + k)[0]
            ids[eng.submit(g, s)] = (g, s)
            submitted += 1
            eng.step()
            collect()
        eng.run()  # drain: the engine goes idle before the next session
        collect()
    assert len(out) == submitted and not ids
    for (g, s), (depth, pred, sd) in out:
        check_query(orc, host, g, s, depth, pred, sd)
    eng.close()
    mg.close()


@pytest.mark.parametrize("lps", [pytest.param(0, id="launch"), pytest.param(2, id="graph2"),
                                 pytest.param(-2, id="persistent2")])
def test_query_ids_recycled_and_pack_released_first(mods, orc, lps):
    """max_queries bounds the queries submitted and not yet collected: ids
    come back after their results are taken, so a long-running engine keeps
    serving.  The engine holds its own reference to the multigraph, which
    may be destroyed first."""
    bfs, batch = mods
    from back40computing_amd import B40CError, _lib
    graphs = [bfs.Graph.rmat(9, 16, 0.57, 0.19, 0.19, 81), bfs.Graph.lattice3d(8, 82)]
    host = [g.to_host() for g in graphs]
    mg = batch.MultiGraph(graphs)
    eng = batch.MultiSourceBFS(mg, label_budget_bytes=1 << 24, max_queries=8)
    set_form(eng, lps)
    nodes = [mg.nodes(0), mg.nodes(1)]
    mg.close()  # the engine keeps the pack alive
    seen = set()
    for rnd in range(4):
        batch_q = [(k % 2, int(orc.pick_sources(host[k % 2][0], 100 * rnd + k, 1)[0])) for k in range(8)]
        ids = {}
        for g, s in batch_q:
            ids[eng.submit(g, s)] = (g, s)
        assert len(ids) == 8
        seen |= set(ids)
        with pytest.raises(B40CError) as ei:
            eng.submit(0, 0)
        assert ei.value.status == _lib.ERR_ADMISSION
        eng.run()
        for q in eng.poll():
            g, s = ids.pop(q)
            depth, pred, sd = eng.result(q, g)  # sizes come from the Python-side copy of the offsets
            assert depth.size == nodes[g]
            check_query(orc, host, g, s, depth, pred, sd)
        assert not ids
    assert seen == set(range(8))
    eng.close()



@pytest.mark.parametrize("per_graph", [pytest.param(2, id="narrow"), pytest.param(12, id="wide")])
@pytest.mark.parametrize("lps", [pytest.param(0, id="launch"), pytest.param(-3, id="persistent3")])
def test_pull_table_overflow(mods, orc, per_graph, lps):
    """More graphs pulling bottom-up in one level than the pull's LDS graph
    table holds (512): 600 copies of one RMAT-8 graph, each given the same
    sources, share every direction decision (a huge alpha sends them
    bottom-up from the first step), so each bottom-up level lists all 600
    in one form (2 queries per graph: the narrow form, 12: the wide
    one) and that form walks every tile of the pack with every graph's bits
    instead.  Every query's labels must still match the oracle."""
    bfs, batch = mods
    g = bfs.Graph.rmat(8, 16, 0.57, 0.19, 0.19, 91)
    off, cols = g.to_host()
    copies = 600
    mg = batch.MultiGraph([g] * copies)
    assert mg.ngraphs == copies
    srcs = [int(x) for x in orc.pick_sources(off, 92, per_graph)]
    ref = {s: orc.bfs(off, cols, s, with_pred=False) for s in srcs}
    eng = batch.MultiSourceBFS(mg, label_budget_bytes=1 << 28)
    set_form(eng, lps)
    eng.set_tuning(alpha=1 << 20)  # m_f * alpha > m_u from the first step: every level after it pulls
    print("pull table overflow: %d graphs x %d queries, levels_per_step=%d" % (copies, per_graph, lps), flush=True)
    res = run_all(eng, mg, [(k, s) for k in range(copies) for s in srcs])
    assert len(res) == copies * per_graph
    for q, ((k, s), (depth, pred, sd)) in res.items():
        rd, _, rsd = ref[s]
        np.testing.assert_array_equal(depth, rd)
        assert orc.check_pred(off, cols, rd, pred) == -1
        assert sd == rsd
    eng.close()
    mg.close()
