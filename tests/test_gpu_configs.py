"""GPU parity at the batched and served configurations' full size.

* Batched: 1,024 queries over 256 RMAT graphs (a=0.57 b=0.19 c=0.19, edge
  factor 16, scale integer-uniform in [10, 18], 4 seeded sources per graph)
  packed into one CSR and run by the multi-source engine: every query's
  depth labels are compared bit-exactly with the serial oracle on its own
  graph, and every predecessor is checked with the reference's invariants
  (test/bfs/test_bfs.cu:387-438).
* Served: resident RMAT-23 plus 512 RMAT graphs of scale 12-16, Poisson
  arrivals (90 % on the big graph) through the continuous-batching server;
  a sample of the served queries is compared with the oracle, and with the
  plain single-query enactor run on the query alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from back40computing_amd import batch, bfs, serve
    return bfs, batch, serve


def test_batched_1024_queries_against_oracle(mods, orc):
    bfs, batch, serve = mods
    graphs, offsets, scales, queries = serve.batched_workload()
    assert len(graphs) == 256 and len(queries) == 1024
    assert min(scales) == 10 and max(scales) == 18
    mg = batch.MultiGraph(graphs)
    budget = int(sum(graphs[g].nodes for g, _ in queries) * 8 * 1.05) + (1 << 20)
    eng = batch.MultiSourceBFS(mg, budget, len(queries) + 8)
    ids = {}
    for g, s in queries:
        ids[eng.submit(g, s)] = (g, s)
    print("batched: %d queries submitted over %d graphs" % (len(ids), len(graphs)), flush=True)
    eng.run()
    done = eng.poll()
    assert sorted(done) == sorted(ids)
    results = {}
    for q in done:
        g, s = ids[q]
        results[(g, s)] = eng.result(q, g)
    eng.close()
    mg.close()
    by_graph = {}
    for g, s in queries:
        by_graph.setdefault(g, []).append(s)
    checked = 0
    for g, srcs in by_graph.items():
        off, cols = graphs[g].to_host()
        refs = orc.bfs_many(off, cols, srcs, threads=16)
        preds = []
        for k, s in enumerate(srcs):
            depth, pred, sd = results[(g, s)]
            np.testing.assert_array_equal(depth, refs[k], err_msg="graph %d source %d" % (g, s))
            assert sd == int(refs[k].max()) + 1
            preds.append(pred)
            checked += 1
        assert orc.check_pred_many(off, cols, refs, preds) == [-1] * len(srcs), "graph %d" % g
    assert checked == 1024
    for g in graphs:
        g.close()


def test_served_rmat23_pool_against_oracle(mods, orc):
    bfs, batch, serve = mods
    from back40computing_amd.sources import pick_sources
    big = bfs.Graph.rmat(23, 16, 0.57, 0.19, 0.19, 7)
    pool_scales = [12 + int(serve._uniform(77, k) * 5) for k in range(512)]
    pool, pool_off = serve.build_pool(pool_scales, seed0=50000)
    assert big.nodes == 1 << 23 and len(pool) == 512
    graphs = [big] + pool
    big_off, big_cols = big.to_host()
    big_src = pick_sources(big_off, 11, 64)
    pool_src = [pick_sources(o, 100 + i, 4) for i, o in enumerate(pool_off)]

    def choose(u, k):
        if u < 0.9:
            return (0, int(big_src[k % len(big_src)]))
        g = 1 + int((u - 0.9) / 0.1 * 512) % 512
        return (g, int(pool_src[g - 1][k % len(pool_src[g - 1])]))

    mg = batch.MultiGraph(graphs)
    cont = serve.ContinuousServer(mg, label_budget_bytes=6 << 30)
    trace = serve.poisson_trace(1500.0, 0.2, choose, seed=31)
    print("served: %d Poisson arrivals over RMAT-23 + 512 pool graphs" % len(trace), flush=True)
    lat, wall = cont.serve(trace)
    assert len(lat) == len(trace) and np.all(lat >= 0)
    # A sample of the trace's queries, run again in one mix and kept: labels
    # against the oracle and against the plain enactor on the query alone.
    sample = []
    for _, q in trace:
        if q not in sample:
            sample.append(q)
        if len(sample) == 24:
            break
    ids = {cont.eng.submit(g, s): (g, s) for g, s in sample}
    cont.eng.run()
    mix = {}
    for q in cont.eng.poll():
        g, s = ids[q]
        mix[(g, s)] = cont.eng.result(q, g)
    assert len(mix) == len(sample)
    plain = serve.PlainServer(graphs, bfs._lib.DIRECTION_OPTIMIZING)
    big_srcs = [s for g, s in sample if g == 0]
    big_refs = orc.bfs_many(big_off, big_cols, big_srcs, threads=16) if big_srcs else []
    big_checks = []
    for (g, s), (depth, pred, sd) in mix.items():
        if g == 0:
            off, cols, ref = big_off, big_cols, big_refs[big_srcs.index(s)]
            big_checks.append((ref, pred))
        else:
            off, cols = pool[g - 1].to_host()
            ref = orc.bfs(off, cols, s, with_pred=False)[0]
            assert orc.check_pred(off, cols, ref, pred) == -1
        np.testing.assert_array_equal(depth, ref, err_msg="graph %d source %d" % (g, s))
        assert sd == int(ref.max()) + 1
        alone, _ = plain.run_one(g, s, want=True)
        np.testing.assert_array_equal(alone, depth)
    # The RMAT-23 queries' predecessors, several at a time on the host.
    assert orc.check_pred_many(big_off, big_cols, [r for r, _ in big_checks],
                               [p for _, p in big_checks]) == [-1] * len(big_checks)
    plain.close()
    cont.eng.close()
    mg.close()
