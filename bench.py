#!/usr/bin/env python3
"""Benchmark: single-source BFS GTEPS on RMAT scale 22 on one MI355X.

Workload: RMAT scale 22 (4,194,304 vertices), edge factor 16, (a,b,c,d) =
(0.57, 0.19, 0.19, 0.05), symmetrized, self loops and duplicates removed,
int32 CSR built on the GPU by radix-sorting packed 64-bit (src,dst) keys.
64 seeded sources of nonzero degree.  One step = Reset + EnactSearch for the
next source (inputs resident in HBM).  GTEPS = sum over steps of
E_trav(src) / elapsed, E_trav = sum of degrees of reached vertices
(test/bfs/test_bfs.cu:359-367).

Prints one JSON line (see the README of the metric in the repo docs).
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


class Hip:
    """Minimal HIP event API over ctypes (timing on the problem's own stream)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def event(self):
        e = ctypes.c_void_p()
        assert self.L.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, e0, e1):
        assert self.L.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert self.L.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        return ms.value

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tuning_for(strategy, graph, arch):
    """Autotuned knobs from the committed tuning table (None: library defaults)."""
    from back40computing_amd import autotune
    return autotune.lookup(arch, strategy, graph.nodes, graph.edges)


def measure_strategy(bfs, b4, hip, graph, srcs, e_trav, v_vis, strategy, steps, warmup, mode=0, tuning=None):
    prob = bfs.CsrProblem()
    b4._lib.check(prob.FromGraph(graph))
    en = bfs.ENACTORS[strategy]()
    if tuning:
        en.set_tuning(**tuning)
    ft = en.GetFrontierType()
    stream = prob.stream()
    lib = b4.lib()

    def step(k, evs=None):
        s = int(srcs[k % len(srcs)])
        b4._lib.check(prob.Reset(ft, 1.3))
        if evs:
            hip.record(evs[0], stream)
        b4._lib.check(lib.b40c_enact_search(en.handle, prob.handle, s, 0, mode))
        if evs:
            hip.record(evs[1], stream)

    for k in range(warmup):
        step(k)
    hip.sync()
    ev = [(hip.event(), hip.event()) for _ in range(steps)]
    t_all0, t_all1 = hip.event(), hip.event()
    hip.sync()
    hip.record(t_all0, stream)
    w0 = time.perf_counter()
    for k in range(steps):
        step(warmup + k, ev[k])
    hip.record(t_all1, stream)
    total_ms = hip.elapsed_ms(t_all0, t_all1)
    hip.sync()
    wall_ms = (time.perf_counter() - w0) * 1e3
    enact_ms = [hip.elapsed_ms(a, b) for a, b in ev]
    idx = [(warmup + k) % len(srcs) for k in range(steps)]
    et = np.array([e_trav[i] for i in idx], np.float64)
    vv = np.array([v_vis[i] for i in idx], np.float64)
    b_alg = 8.0 * et + 16.0 * vv
    res = {
        "gteps": float(et.sum() / (total_ms * 1e-3) / 1e9),
        "gteps_enact_only": float(et.sum() / (np.sum(enact_ms) * 1e-3) / 1e9),
        "ms_per_step": total_ms / steps,
        "ms_per_search_enact": float(np.mean(enact_ms)),
        "ms_per_search_p50": float(np.median(enact_ms)),
        "host_wall_ms_per_step": wall_ms / steps,
        "achieved_gbs": float(b_alg.sum() / (np.sum(enact_ms) * 1e-3) / 1e9),
        "b_alg_per_search": float(b_alg.mean()),
    }
    prob.close()
    en.close()
    return res


def pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def bench_lattice(bfs, b4, hip, info, width=200, sources=4, steps=8, only=None):
    """High-diameter workload: 200^3 6-neighbour lattice, ids permuted from
    seed 0, `sources` seeded sources; GTEPS per strategy (one launch per
    level, levels captured in hipGraphs where the strategy allows)."""
    from back40computing_amd.sources import pick_sources
    g = bfs.Graph.lattice3d(width, 0)
    off = np.empty(g.nodes + 1, np.int32)
    b4._lib.check(b4.lib().b40c_graph_copy_to_host(g.handle, b4._lib.ptr(off), None))
    deg = np.diff(off).astype(np.int64)
    srcs = pick_sources(off, 3, sources)
    out = {"workload": "lattice%d^3_perm_seed0" % width, "vertices": g.nodes, "csr_edges": g.edges,
           "sources": len(srcs)}
    for s_id, mode, name in ((b4.EXPAND_CONTRACT, b4.MODE_ITERATIVE, "expand_contract_iterative"),
                             (b4.EXPAND_CONTRACT, b4.MODE_FUSED, "expand_contract_persistent"),
                             (b4.CONTRACT_EXPAND, b4.MODE_FUSED, "contract_expand_persistent"),
                             (b4.TWO_PHASE, b4.MODE_FUSED, "two_phase_persistent"),
                             (b4.HYBRID, b4.MODE_DEFAULT, "hybrid"),
                             (b4.DIRECTION_OPTIMIZING, b4.MODE_DEFAULT, "direction_optimizing")):
        if only and name not in only:
            continue
        t = tuning_for(s_id, g, info["arch"])
        e_trav, v_vis, levels = [], [], []
        for s in srcs:
            depth, _, st = bfs.bfs(g, int(s), strategy=s_id, mode=mode, tuning=t)
            m = depth >= 0
            e_trav.append(int(deg[m].sum()))
            v_vis.append(int(m.sum()))
            levels.append(st["search_depth"])
        r = measure_strategy(bfs, b4, hip, g, srcs, e_trav, v_vis, s_id, steps, 1, mode=mode, tuning=t)
        out[name] = {"gteps": round(r["gteps"], 3), "ms_per_search": round(r["ms_per_search_enact"], 3),
                     "levels_mean": round(float(np.mean(levels)), 1),
                     "us_per_level": round(1e3 * r["ms_per_search_enact"] / max(1.0, float(np.mean(levels))), 2),
                     "roofline_frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4)}
        log("lattice %s: %s" % (name, out[name]))
    g.close()
    return out


BATCH_LEVELS_PER_STEP = 4


def bench_batched(b4, info):
    """1,024 queries over 256 RMAT graphs (scale integer-uniform in [10,18],
    4 sources each): the batched multi-source engine vs the plain
    single-query enactor run query by query, labels compared for every query."""
    from back40computing_amd import serve
    from back40computing_amd.batch import MultiGraph, MultiSourceBFS
    from back40computing_amd import autotune
    t0 = time.perf_counter()
    graphs, offsets, scales, queries = serve.batched_workload()
    build_s = time.perf_counter() - t0
    mg = MultiGraph(graphs)
    log("batched workload: %d graphs, %d queries, V=%d E=%d (built in %.2f s)" % (
        len(graphs), len(queries), int(mg.vbase[-1]), int(mg.ebase[-1]), build_s))
    plain = serve.PlainServer(graphs, b4.DIRECTION_OPTIMIZING, tuning_for=lambda g: autotune.lookup(
        info["arch"], b4.DIRECTION_OPTIMIZING, graphs[g].nodes, graphs[g].edges))
    for g, s in queries[:16]:
        plain.run_one(g, s)
    plain_ms = []
    t0 = time.perf_counter()
    for g, s in queries:
        t1 = time.perf_counter()
        plain.run_one(g, s)
        plain_ms.append((time.perf_counter() - t1) * 1e3)
    plain_s = time.perf_counter() - t0
    budget = int(sum(graphs[g].nodes for g, _ in queries) * 8 * 1.05) + (1 << 20)

    def run_batched(levels_per_step, persistent=False):
        eng = MultiSourceBFS(mg, budget, len(queries) + 8)
        eng.set_levels_per_step(levels_per_step)
        eng.set_persistent(persistent)
        ids, done_ms = {}, {}
        t0 = time.perf_counter()
        for g, s in queries:
            ids[eng.submit(g, s)] = (g, s)
        steps = 0
        while True:
            r, w = eng.step()
            steps += 1
            now = (time.perf_counter() - t0) * 1e3
            for q in eng.poll():
                done_ms[q] = now
            if r == 0 and w == 0:
                break
        return eng, ids, done_ms, steps, time.perf_counter() - t0

    # Three launch forms of the same engine: five launches and a host
    # synchronisation per level; a captured hipGraph replaying
    # BATCH_LEVELS_PER_STEP levels per synchronisation; one persistent launch
    # of BATCH_LEVELS_PER_STEP levels (grid barriers between the steps).
    forms = {}
    for name, lps, pers in (("launch", 0, False), ("graph", BATCH_LEVELS_PER_STEP, False),
                            ("persistent", BATCH_LEVELS_PER_STEP, True)):
        run_batched(lps, pers)[0].close()  # warm-up (capture, first-touch of the label arena)
        eng, ids, done_ms, steps, batch_s = run_batched(lps, pers)
        mismatches = 0
        for q, (g, s) in ids.items():
            d_b, _, _ = eng.result(q, g, want_pred=False)
            d_p, _ = plain.run_one(g, s, want=True)
            mismatches += int(not np.array_equal(d_b, d_p))
        eng.close()
        lat = list(done_ms.values())
        forms[name] = {"levels_per_step": lps, "persistent": pers, "steps": steps, "batched_qps": round(len(queries) / batch_s, 1),
                       "speedup": round(plain_s / batch_s, 2), "batched_ms_p50": round(pct(lat, 50), 3),
                       "batched_ms_p99": round(pct(lat, 99), 3), "label_mismatches": mismatches}
        log("batched %s: %s" % (name, forms[name]))
    best = max(forms.values(), key=lambda f: f["batched_qps"])
    # The same arrival for the plain enactor: all queries submitted at once
    # and served one after another, so a query finishes when every query
    # before it and itself have run.
    plain_done = np.cumsum(plain_ms).tolist()
    mismatches = sum(f["label_mismatches"] for f in forms.values())
    res = {
        "queries": len(queries), "graphs": len(graphs), "levels_per_step": best["levels_per_step"],
        "persistent": best["persistent"],
        "plain_qps": round(len(queries) / plain_s, 1), "batched_qps": best["batched_qps"],
        "speedup": best["speedup"],
        "plain_ms_p50": round(pct(plain_ms, 50), 4), "plain_ms_p99": round(pct(plain_ms, 99), 4),
        "batched_ms_p50": best["batched_ms_p50"], "batched_ms_p99": best["batched_ms_p99"],
        "plain_fifo_ms_p50": round(pct(plain_done, 50), 3), "plain_fifo_ms_p99": round(pct(plain_done, 99), 3),
        "forms": forms,
        "labels_identical": mismatches == 0, "label_mismatches": mismatches,
        "note": "batched latency = batch start -> query finished (all 1,024 submitted at once); plain_ms = one "
                "query alone (Reset + EnactSearch, direction-optimizing); plain_fifo_ms = the same all-at-once "
                "arrival served one query after another by the plain enactor; headline fields from the fastest "
                "launch form, labels of every form compared with the plain enactor's for every query",
    }
    plain.close()
    mg.close()
    log("batched:", res)
    return res


def bench_served(b4, info, duration_s=1.5):
    """Resident RMAT-23 + 512 RMAT graphs of scale 12-16; Poisson arrivals
    (90 % on the big graph); continuous batching vs a FIFO plain server."""
    from back40computing_amd import serve
    from back40computing_amd.batch import MultiGraph
    from back40computing_amd.bfs import Graph
    from back40computing_amd import autotune
    t0 = time.perf_counter()
    big = Graph.rmat(23, 16, 0.57, 0.19, 0.19, 7)
    pool_scales = [12 + int(serve._uniform(77, k) * 5) for k in range(512)]
    pool, pool_off = serve.build_pool(pool_scales, seed0=50000)
    graphs = [big] + pool
    big_off = np.empty(big.nodes + 1, np.int32)
    b4._lib.check(b4.lib().b40c_graph_copy_to_host(big.handle, b4._lib.ptr(big_off), None))
    big_src = pick_sources_cached(big_off, 11, 4096)
    pool_src = [pick_sources_cached(o, 100 + i, 8) for i, o in enumerate(pool_off)]
    mg = MultiGraph(graphs)
    log("served workload: V=%d E=%d (built in %.2f s)" % (int(mg.vbase[-1]), int(mg.ebase[-1]),
                                                          time.perf_counter() - t0))

    def choose(u, k):
        if u < 0.9:
            return (0, int(big_src[k % len(big_src)]))
        g = 1 + int((u - 0.9) / 0.1 * 512) % 512
        return (g, int(pool_src[g - 1][k % len(pool_src[g - 1])]))

    plain = serve.PlainServer(graphs, b4.DIRECTION_OPTIMIZING, tuning_for=lambda g: autotune.lookup(
        info["arch"], b4.DIRECTION_OPTIMIZING, graphs[g].nodes, graphs[g].edges))
    probe = [choose(serve._uniform(5, k), k) for k in range(40)]
    for g, s in probe[:8]:
        plain.run_one(g, s)
    t1 = time.perf_counter()
    for g, s in probe:
        plain.run_one(g, s)
    service = (time.perf_counter() - t1) / len(probe)
    cap = 1.0 / service
    out = {"plain_service_ms": round(service * 1e3, 3), "plain_capacity_qps": round(cap, 1), "points": []}
    rate_plain = 0.7 * cap
    trace = serve.poisson_trace(rate_plain, duration_s, choose, seed=21)
    log("served: plain service %.3f ms, capacity %.0f qps; plain trace of %d queries" % (
        service * 1e3, cap, len(trace)))
    lat, wall = plain.serve(trace, log=log)
    out["points"].append({"server": "plain", "offered_qps": round(rate_plain, 1), "queries": len(trace),
                          "achieved_qps": round(len(trace) / wall, 1), "p50_ms": round(pct(lat, 50) * 1e3, 3),
                          "p99_ms": round(pct(lat, 99) * 1e3, 3)})
    # Continuous batching in both launch forms: one level per step with five
    # launches, and one captured single-level hipGraph per step.
    for lps in (0, 1):
        cont = serve.ContinuousServer(mg, label_budget_bytes=6 << 30, log=log, levels_per_step=lps)
        for mult in (2, 3, 4, 5, 10):
            rate = mult * rate_plain
            trace = serve.poisson_trace(rate, duration_s, choose, seed=21 + mult)
            log("served: continuous (levels_per_step %d) trace of %d queries at %.0f qps" % (lps, len(trace), rate))
            lat, wall = cont.serve(trace)
            achieved = len(trace) / wall
            out["points"].append({"server": "continuous", "levels_per_step": lps, "offered_qps": round(rate, 1),
                                  "queries": len(trace), "achieved_qps": round(achieved, 1),
                                  "p50_ms": round(pct(lat, 50) * 1e3, 3), "p99_ms": round(pct(lat, 99) * 1e3, 3)})
            log("served: point", out["points"][-1])
            if achieved < 0.8 * rate:
                break  # saturated: higher offered rates only lengthen the queue
        if lps == 0:
            cont.eng.close()
    # Alone vs in the mix: labels of a sample of queries served both ways.
    sample = [choose(serve._uniform(99, k), k) for k in range(24)]
    ids = {cont.eng.submit(g, s): (g, s) for g, s in sample}
    cont.eng.run()
    same = 0
    for q in cont.eng.poll():
        g, s = ids[q]
        d_mix, _, _ = cont.eng.result(q, g, want_pred=False)
        d_alone, _ = plain.run_one(g, s, want=True)
        same += int(np.array_equal(d_mix, d_alone))
    out["labels_identical"] = same == len(sample)
    p_plain = out["points"][0]
    best = max((p for p in out["points"][1:] if p["p99_ms"] <= p_plain["p99_ms"]),
               key=lambda p: p["achieved_qps"], default=None)
    out["qps_ratio_at_equal_or_lower_p99"] = (round(best["achieved_qps"] / p_plain["achieved_qps"], 2)
                                              if best else None)
    cont.eng.close()
    plain.close()
    mg.close()
    log("served:", out)
    return out


def bench_cohorts(b4, info, ks=(1, 4, 16, 64), reps=3):
    """Cost of one cohort of k queries on the served pack's RMAT-23 graph
    (RMAT-23 + 512 RMAT graphs of scale 12-16): all k admitted at one level
    boundary and run to completion by the batched engine, one level per
    step; beside it the plain direction-optimizing enactor per query.  The
    served analysis and its kernel trace rest on this leg."""
    from back40computing_amd import serve
    from back40computing_amd.batch import MultiGraph, MultiSourceBFS
    from back40computing_amd.bfs import Graph
    from back40computing_amd import autotune
    big = Graph.rmat(23, 16, 0.57, 0.19, 0.19, 7)
    pool_scales = [12 + int(serve._uniform(77, k) * 5) for k in range(512)]
    pool, _ = serve.build_pool(pool_scales, seed0=50000)
    mg = MultiGraph([big] + pool)
    off = np.empty(big.nodes + 1, np.int32)
    b4._lib.check(b4.lib().b40c_graph_copy_to_host(big.handle, b4._lib.ptr(off), None))
    srcs = [int(s) for s in pick_sources_cached(off, 11, max(ks) * reps)]
    eng = MultiSourceBFS(mg, 6 << 30)
    out = {"workload": "rmat23_served_pack_cohorts", "cohorts": []}
    for k in ks:
        for rep in range(reps):
            for s in srcs[rep * k:(rep + 1) * k]:
                eng.submit(0, s)
            t0 = time.perf_counter()
            levels = 0
            while True:
                r, w = eng.step()
                levels += 1
                if r == 0 and w == 0:
                    break
            ms = (time.perf_counter() - t0) * 1e3
            for q in eng.poll():
                eng.release(q)
            out["cohorts"].append({"k": k, "levels": levels, "ms": round(ms, 3), "ms_per_query": round(ms / k, 4)})
            log("cohort k=%d: %.3f ms (%.4f ms per query, %d levels)" % (k, ms, ms / k, levels))
    plain = serve.PlainServer([big], b4.DIRECTION_OPTIMIZING, tuning_for=lambda g: autotune.lookup(
        info["arch"], b4.DIRECTION_OPTIMIZING, big.nodes, big.edges))
    for s in srcs[:4]:
        plain.run_one(0, s)
    t0 = time.perf_counter()
    for s in srcs[:32]:
        plain.run_one(0, s)
    out["plain_ms_per_query"] = round((time.perf_counter() - t0) * 1e3 / 32, 4)
    log("cohorts: plain %.4f ms per query" % out["plain_ms_per_query"])
    # Plain searches side by side: N host threads, each with its own
    # problem, enactor and HIP stream (the library's calls release the GIL).
    import threading
    from back40computing_amd import bfs as bfsmod
    t = autotune.lookup(info["arch"], b4.DIRECTION_OPTIMIZING, big.nodes, big.edges)
    out["plain_concurrent"] = []
    for n in (1, 2, 4, 8):
        slots = []
        for k in range(n):
            p = bfsmod.CsrProblem()
            b4._lib.check(p.FromGraph(big))
            en = bfsmod.ENACTORS[b4.DIRECTION_OPTIMIZING]()
            if t:
                en.set_tuning(**t)
            b4._lib.check(p.Reset(en.GetFrontierType(), 1.3))
            b4._lib.check(b4.lib().b40c_enact_search(en.handle, p.handle, srcs[k], 0, 0))
            slots.append((p, en))
        per = 24

        def run(k):
            p, en = slots[k]
            for i in range(per):
                b4._lib.check(p.Reset(en.GetFrontierType(), 1.3))
                b4._lib.check(b4.lib().b40c_enact_search(en.handle, p.handle, srcs[(k * per + i) % len(srcs)], 0, 0))

        t0 = time.perf_counter()
        th = [threading.Thread(target=run, args=(k,)) for k in range(n)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        rate = n * per / (time.perf_counter() - t0)
        out["plain_concurrent"].append({"streams": n, "searches_per_s": round(rate, 1)})
        log("cohorts: %d plain searches side by side: %.0f searches/s" % (n, rate))
        for p, en in slots:
            p.close()
            en.close()
    eng.close()
    plain.close()
    mg.close()
    return out


def _ingest_one(graph_io, path, fmt, off, cols, write_s=None, reps=2):
    """Best of `reps` reads of one file: host parse (b40c_edge_list_read)
    and device CSR build (b40c_graph_from_edge_list); the CSR must be the
    graph the file was written from."""
    size = os.path.getsize(path)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        el = graph_io.EdgeList(path, fmt)
        t1 = time.perf_counter()
        g = el.to_graph()
        t2 = time.perf_counter()
        if best is None or t2 - t0 < best[0] + best[1]:
            best = (t1 - t0, t2 - t1, el.edges)
        el.close()
    o2, c2 = g.to_host()
    parse_s, build_s, m = best
    out = {"file_mb": round(size / 1e6, 1), "edges": m, "parse_s": round(parse_s, 4), "build_s": round(build_s, 4),
           "parse_mb_per_s": round(size / 1e6 / parse_s, 1),
           "medges_per_s": round(m / (parse_s + build_s) / 1e6, 1),
           "csr_identical": bool(np.array_equal(o2, off) and np.array_equal(c2, cols))}
    if write_s is not None:
        out["write_s"] = round(write_s, 4)
    g.close()
    return out


def bench_ingest(scale=20, text_scale=16):
    """Graph files (MatrixMarket, METIS, DIMACS; test_bfs.cu:976-1037): an
    RMAT-`scale` graph written as METIS by the library's writer, an
    RMAT-`text_scale` graph written as MatrixMarket and DIMACS text (one
    directed entry per CSR edge, rows in order), each read back: parse on the
    host cores, device CSR built on the GPU, checked identical to the
    graph written."""
    import tempfile
    from back40computing_amd import graph_io
    from back40computing_amd.bfs import Graph
    out = {"threads": min(16, os.cpu_count() or 1)}
    with tempfile.TemporaryDirectory() as d:
        g = Graph.rmat(scale, 16, 0.57, 0.19, 0.19, 3)
        off, cols = g.to_host()
        path = os.path.join(d, "rmat.graph")
        t = time.perf_counter()
        graph_io.write_metis(g, path)
        out["metis_rmat%d" % scale] = _ingest_one(graph_io, path, "metis", off, cols, time.perf_counter() - t)
        g.close()
        os.remove(path)
        g = Graph.rmat(text_scale, 16, 0.57, 0.19, 0.19, 4)
        off, cols = g.to_host()
        rows = (np.repeat(np.arange(g.nodes), np.diff(off)) + 1).tolist()
        c1 = (cols.astype(np.int64) + 1).tolist()
        mpath = os.path.join(d, "rmat.mtx")
        with open(mpath, "w") as f:
            f.write("%%%%MatrixMarket matrix coordinate pattern general\n%d %d %d\n" % (g.nodes, g.nodes, g.edges))
            f.write("\n".join("%d %d" % rc for rc in zip(c1, rows)))
            f.write("\n")
        out["market_rmat%d" % text_scale] = _ingest_one(graph_io, mpath, "market", off, cols)
        dpath = os.path.join(d, "rmat.gr")
        with open(dpath, "w") as f:
            f.write("p sp %d %d\n" % (g.nodes, g.edges))
            f.write("\n".join("a %d %d 1" % rc for rc in zip(rows, c1)))
            f.write("\n")
        out["dimacs_rmat%d" % text_scale] = _ingest_one(graph_io, dpath, "dimacs", off, cols)
        g.close()
    return out


def bench_microbench(bfs, graph, srcs, strategy, tuning, nsrc=4):
    """Per-stage microbenchmark search (status lookup and neighbour gather,
    each as a bench pass and a real pass, HIP events per launch) on the
    headline graph from its first sources, both modes; the second search of
    each source is the one counted.  Stage totals carry algorithmic bytes
    (back40computing_amd.microbench) and GB/s.  The labels are compared with
    the traversal enactor's."""
    from back40computing_amd import microbench
    out = {}
    for mode in ("gather_lookup", "cull"):
        levels_all = []
        match = True
        for s in srcs[:nsrc]:
            microbench.search(graph, int(s), mode, want_depth=False)
            depth, levels = microbench.search(graph, int(s), mode)
            ref, _, _ = bfs.bfs(graph, int(s), strategy=strategy, tuning=tuning)
            match = match and bool(np.array_equal(depth, ref))
            levels_all += levels
        summ = microbench.summarize(levels_all)
        out[mode] = {k: summ[k] for k in ("lookup", "bench_lookup", "offsets", "gather", "bench_gather")}
        out[mode].update({k: summ[k] for k in summ if k.endswith("_frac")})
        out[mode]["edges_gathered"] = summ["edges_gathered"]
        out[mode]["labels_match_enactor"] = match
    out["sources"] = len(srcs[:nsrc])
    return out


def bench_primitives(hip, n_scan=1 << 28, n_trim=1 << 27, reps=5):
    """Consecutive removal and prefix scan at HBM-filling sizes, timed with
    HIP events on the stream they run on; algorithmic bytes: scan 4 B read +
    4 B written per element, trim 8 B read per key + 8 B written per kept
    key.  Results are checked against numpy on the host."""
    from back40computing_amd import _lib, primitives
    from back40computing_amd.radix_sort import DeviceArray, DoubleBuffer
    out = {}
    rng = np.random.default_rng(1)
    x = rng.integers(0, 1 << 20, n_scan).astype(np.uint32)
    d = DeviceArray(x.nbytes)
    d.upload(x)
    e0, e1 = hip.event(), hip.event()
    best = None
    for _ in range(reps):
        d.upload(x)
        hip.sync()
        hip.record(e0, None)
        primitives.Scan().Scan(d, d, n_scan, True)
        hip.record(e1, None)
        ms = hip.elapsed_ms(e0, e1)
        best = ms if best is None else min(best, ms)
    got = d.download(np.uint32, n_scan)
    ref = (np.cumsum(x, dtype=np.uint64) - x).astype(np.uint32)
    gbs = 8.0 * n_scan / (best * 1e-3) / 1e9
    out["scan_u32"] = {"n": n_scan, "ms": round(best, 4), "gb_per_s": round(gbs, 1),
                       "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "correct": bool(np.array_equal(got, ref))}
    d.free()
    keys = np.cumsum(rng.integers(0, 2, n_trim)).astype(np.uint64)  # non-decreasing, runs of equal keys
    buf = DoubleBuffer(np.uint64, n_trim)
    best, m = None, 0
    for _ in range(reps):
        buf.selector = 0
        buf.d_keys[0].upload(keys)
        hip.sync()
        hip.record(e0, None)
        m = primitives.ConsecutiveRemoval().Trim(buf, n_trim)
        hip.record(e1, None)
        ms = hip.elapsed_ms(e0, e1)
        best = ms if best is None else min(best, ms)
    head = np.ones(n_trim, bool)
    head[1:] = keys[1:] != keys[:-1]
    ok = m == int(head.sum()) and np.array_equal(buf.d_keys[buf.selector].download(np.uint64, m), keys[head])
    gbs = (8.0 * n_trim + 8.0 * m) / (best * 1e-3) / 1e9
    out["trim_u64"] = {"n": n_trim, "kept": m, "ms": round(best, 4), "gb_per_s": round(gbs, 1),
                       "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "correct": bool(ok)}
    for a in buf.d_keys:
        a.free()
    return out


def bench_sort(hip, n=1 << 27, reps=3):
    """GPU radix sort at the CSR build's size (test/radix_sort/test_sort.cu:
    155-163 reports 10^9 keys/s): n random 64-bit keys (8 digit places), and
    n uint32 keys with uint32 values (4 places), through a sort enactor that
    keeps its scratch.  Checked by size-independent properties: the output
    is non-decreasing and has the input's key sum and xor (and, for pairs,
    every value still travels with its key).  Algorithmic bytes: a key read
    and written once per digit place (16 B per 64-bit key per place, 16 B per
    key-value pair per place); this design moves 24 B per 64-bit key per
    place (upsweep read, downsweep read and write)."""
    from back40computing_amd import radix_sort as rs
    rng = np.random.default_rng(5)
    out = {}
    en = rs.Enactor()
    for name, kdt, with_vals, passes, alg_b in (("u64_keys", np.uint64, False, 8, 16.0),
                                                 ("u32_pairs", np.uint32, True, 4, 16.0)):
        info = np.iinfo(kdt)
        keys = rng.integers(0, info.max, n, dtype=kdt, endpoint=True)
        db = rs.DoubleBuffer(kdt, n, with_values=with_vals)
        vals = None
        if with_vals:
            vals = (keys * np.uint32(2654435761)).astype(np.uint32)  # a function of the key, checkable after sort
        best = None
        for _ in range(reps):
            db.selector = 0
            db.d_keys[0].upload(keys)
            if with_vals:
                db.d_values[0].upload(vals)
            hip.sync()
            t0 = time.perf_counter()
            st = en.Sort(db, n)
            dt = time.perf_counter() - t0
            assert st == 0
            best = dt if best is None else min(best, dt)
        got = db.d_keys[db.selector].download(kdt, n)
        ok = bool(np.all(got[1:] >= got[:-1]))
        ok = ok and int(got.sum(dtype=np.uint64)) == int(keys.sum(dtype=np.uint64))
        ok = ok and int(np.bitwise_xor.reduce(got)) == int(np.bitwise_xor.reduce(keys))
        if with_vals:
            gv = db.d_values[db.selector].download(np.uint32, n)
            ok = ok and bool(np.array_equal(gv, (got * np.uint32(2654435761)).astype(np.uint32)))
        gbs = alg_b * passes * n / best / 1e9
        out[name] = {"n": n, "passes": passes, "ms": round(best * 1e3, 3), "gkeys_per_s": round(n / best / 1e9, 3),
                     "alg_gb_per_s": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "correct": ok}
        log("sort %s: %s" % (name, out[name]))
        del got, keys
        for a in db.d_keys + (db.d_values or []):
            a.free()
    en.close()
    out["note"] = ("host-timed synchronous sorts (selector read back), best of %d; alg bytes 16 B per key per "
                   "digit place" % reps)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline_leg(b4, bfs, strategy, tuning, graph22, off22, srcs22, scale22, budget_s=10.0, sources22=6):
    """The reference's only CPU path, SimpleReferenceBfs (test/bfs/test_bfs.cu:
    507-578), restated serially in oracle/oracle.cpp and timed on one host
    core.  The designated measurement is the RMAT scale-16 graph: its 64
    seeded sources are run round after round until `budget_s` of wall time,
    with steady_clock wall time and getrusage user time per run (the
    reference's CpuTimer, test/b40c_test_util.h:503-522).  The same oracle on
    `sources22` sources of the RMAT-22 headline graph is reported beside it.
    The GPU labels of every timed source are compared with the oracle's."""
    import oracle
    from back40computing_amd.sources import pick_sources
    off16, cols16 = oracle.build_rmat_csr(16, 16, 0.57, 0.19, 0.19, 0)
    g16 = bfs.Graph.rmat(16, 16, 0.57, 0.19, 0.19, 0)
    o16, c16 = g16.to_host()
    csr_same = bool(np.array_equal(o16, off16) and np.array_equal(c16, cols16))
    srcs16 = [int(s) for s in pick_sources(off16, 1, 64)]
    deg16 = np.diff(off16).astype(np.int64)
    same16 = 0
    tot_e = tot_wall = tot_user = 0.0
    runs = 0
    t_start = time.perf_counter()
    first = True
    while first or time.perf_counter() - t_start < budget_s:
        for s in srcs16:
            d, sd, wall_ms, user_ms = oracle.bfs_timed(off16, cols16, s)
            tot_e += float(deg16[d >= 0].sum())
            tot_wall += wall_ms
            tot_user += user_ms
            runs += 1
            if first:
                d_gpu, _, st = bfs.bfs(g16, s, strategy=strategy)
                same16 += int(np.array_equal(d_gpu, d) and st["search_depth"] == sd)
        first = False
    g16.close()
    out = {"value": round(tot_e / (tot_wall * 1e-3) / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
           "sample": "serial FIFO BFS (SimpleReferenceBfs restated) on RMAT-16 (a=0.57 b=0.19 c=0.19, ef 16, "
                     "sym, dedup, V=%d E=%d), its 64 seeded sources repeated to %d runs (%.1f s), steady_clock "
                     "wall time, 1 thread" % (off16.size - 1, cols16.size, runs, tot_wall * 1e-3),
           "user_time_gteps": round(tot_e / (max(tot_user, 1e-9) * 1e-3) / 1e9, 5),
           "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "gpu_csr_identical": csr_same, "gpu_labels_identical": "%d/%d" % (same16, len(srcs16))}
    # Secondary: the headline graph.
    if sources22 > 0:
        E = graph22.edges
        cols22 = np.empty(E, np.int32)
        b4._lib.check(b4.lib().b40c_graph_copy_to_host(graph22.handle, None, b4._lib.ptr(cols22)))
        deg22 = np.diff(off22).astype(np.int64)
        te = tw = tu = 0.0
        same22 = 0
        for s in srcs22[:sources22]:
            d, sd, wall_ms, user_ms = oracle.bfs_timed(off22, cols22, int(s))
            te += float(deg22[d >= 0].sum())
            tw += wall_ms
            tu += user_ms
            d_gpu, _, st = bfs.bfs(graph22, int(s), strategy=strategy, tuning=tuning)
            same22 += int(np.array_equal(d_gpu, d) and st["search_depth"] == sd)
        out["rmat%d" % scale22] = {"gteps_wall": round(te / (tw * 1e-3) / 1e9, 5),
                                   "gteps_user": round(te / (max(tu, 1e-9) * 1e-3) / 1e9, 5),
                                   "sources": min(sources22, len(srcs22)),
                                   "gpu_labels_identical": "%d/%d" % (same22, min(sources22, len(srcs22)))}
    return out


def pick_sources_cached(off, seed, count):
    from back40computing_amd.sources import pick_sources
    return pick_sources(off, seed, count)


def probe(args):
    """Short run for hardware-counter passes (rocprofv3 --pmc): a 1 GiB
    device-to-device copy, whose FETCH_SIZE / WRITE_SIZE readings scale the
    counters to bytes, then `--steps` headline searches (Reset + EnactSearch)
    on the RMAT graph with the tuned knobs."""
    import back40computing_amd as b4
    from back40computing_amd import bfs
    from back40computing_amd.sources import pick_sources
    hip = Hip()
    hip.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    nbytes = 1 << 30
    bufs = [ctypes.c_void_p(), ctypes.c_void_p()]
    for bptr in bufs:
        b4._lib.check(b4.lib().b40c_device_alloc(ctypes.byref(bptr), nbytes))
    assert hip.L.hipMemcpy(bufs[1], bufs[0], nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    hip.sync()
    log("probe: copied %d bytes device to device" % nbytes)
    for bptr in bufs:
        b4._lib.check(b4.lib().b40c_device_free(bptr))
    info = bfs.device_info()
    graph = bfs.Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, 0)
    off = np.empty(graph.nodes + 1, np.int32)
    b4._lib.check(b4.lib().b40c_graph_copy_to_host(graph.handle, b4._lib.ptr(off), None))
    srcs = pick_sources(off, 1, args.sources)
    prob = bfs.CsrProblem()
    b4._lib.check(prob.FromGraph(graph))
    en = bfs.ENACTORS[args.strategy]()
    t = tuning_for(args.strategy, graph, info["arch"])
    if t:
        en.set_tuning(**t)
    for k in range(args.steps):
        s = int(srcs[k % len(srcs)])
        b4._lib.check(prob.Reset(en.GetFrontierType(), 1.3))
        b4._lib.check(b4.lib().b40c_enact_search(en.handle, prob.handle, s, 0, b4.MODE_DEFAULT))
        hip.sync()
        log("probe: search %d from %d done" % (k, s))
    prob.close()
    en.close()


def measured_traffic(workload, strategy):
    """HBM bytes per search from the committed counter summary of the same
    workload and strategy (two rocprofv3 --pmc passes, FETCH_SIZE and
    WRITE_SIZE, each scaled by its reading for a copy of known size), and the
    L2 hit rate of its search kernels (a TCC_HIT_sum / TCC_MISS_sum pass);
    traffic is null when there is no summary for this configuration."""
    path = os.path.join(ROOT, "profiles", "rmat22_direction_optimizing_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {"traffic": None}
    if t.get("workload") != workload or t.get("strategy") != strategy:
        return {"traffic": None}
    out = {"traffic": t["traffic_bytes_per_search"], "traffic_unit": "bytes per search",
           "traffic_source": "profiles/rmat22_direction_optimizing_traffic.json (commit %s)" % t.get("commit")}
    if "l2_hit_rate" in t:
        out["l2_hit_rate"] = t["l2_hit_rate"]
    return out


def roofline_block(achieved_gbs, ms_per_search, workload, strategy):
    """achieved/frac: algorithmic bytes (B_alg = 8 E_trav + 16 V_vis per
    search) per second, so for a direction-optimizing search they also count
    the traffic its bottom-up steps avoid.  traffic_gbs/traffic_frac: the
    HBM bytes a search really moves (rocprofv3 counters, committed summary)
    over the same measured search time."""
    out = {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "achieved_kind": "algorithmic bytes (B_alg) per second"}
    out.update(measured_traffic(workload, strategy))
    if out.get("traffic"):
        tg = out["traffic"] / (ms_per_search * 1e-3) / 1e9
        out["traffic_gbs"] = round(tg, 1)
        out["traffic_frac"] = round(tg / HBM_PEAK_GBS, 4)
    return out


def run_child(which, timeout_s=420):
    """Runs `bench.py --only <which>` in a child process (its progress goes to
    our stderr) and returns its result dict, or an error record."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--only", which]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        log("%s workload: timed out after %d s" % (which, timeout_s))
        return {"error": "timed out after %d s" % timeout_s}
    for line in reversed(r.stdout.splitlines()):
        if line.startswith("RESULT "):
            return json.loads(line[len("RESULT "):])
    log("%s workload: child exited with status %d and no result" % (which, r.returncode))
    return {"error": "child exited with status %d" % r.returncode}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--strategy", type=int, default=5, help="headline strategy (5 = direction-optimizing)")
    ap.add_argument("--compare-steps", type=int, default=16, help="steps for the other strategies (0: skip)")
    ap.add_argument("--cpu-sources", type=int, default=6,
                    help="RMAT-22 oracle BFS sources timed beside the RMAT-16 cpu_baseline (0: skip the CPU baseline)")
    ap.add_argument("--no-sort", action="store_true", help="skip the radix-sort workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched 1,024-query workload")
    ap.add_argument("--no-served", action="store_true", help="skip the served Poisson workload")
    ap.add_argument("--no-lattice", action="store_true", help="skip the 200^3 lattice (high-diameter) workload")
    ap.add_argument("--no-ingest", action="store_true",
                    help="skip the graph-file read/build, primitives and per-stage microbenchmark workloads")
    ap.add_argument("--only", choices=["batched", "served", "lattice", "sort", "cohorts"],
                    help="run one secondary workload alone and print its result (profiling runs)")
    ap.add_argument("--lattice-strategies", default=None,
                    help="with --only lattice: comma-separated subset (e.g. expand_contract_persistent)")
    ap.add_argument("--probe", action="store_true",
                    help="counter-collection run: a 1 GiB copy plus --steps searches, no JSON line")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box here has a single MI355X, and the "
              "multi-GPU partitioned BFS is out of scope (--gpus %d refused)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.probe:
        probe(args)
        return
    if args.only:
        import back40computing_amd as b4
        from back40computing_amd import bfs
        info = bfs.device_info()
        if args.only == "batched":
            res = bench_batched(b4, info)
        elif args.only == "served":
            res = bench_served(b4, info)
        elif args.only == "cohorts":
            res = bench_cohorts(b4, info)
        elif args.only == "lattice":
            res = bench_lattice(bfs, b4, Hip(), info,
                                only=args.lattice_strategies.split(",") if args.lattice_strategies else None)
        else:
            res = bench_sort(Hip())
        print("RESULT " + json.dumps(res), flush=True)
        return

    import back40computing_amd as b4
    from back40computing_amd import bfs
    from back40computing_amd.sources import pick_sources

    hip = Hip()
    info = bfs.device_info()
    log("device", info)

    # ---- graph build on the GPU (timed separately; not part of a step) ----
    hip.sync()
    t0 = time.perf_counter()
    graph = bfs.Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, 0)
    hip.sync()
    build_s = time.perf_counter() - t0
    n, E = graph.nodes, graph.edges
    log("rmat-%d built on GPU: V=%d E=%d in %.3f s" % (args.scale, n, E, build_s))
    off = np.empty(n + 1, np.int32)
    b4._lib.check(b4.lib().b40c_graph_copy_to_host(graph.handle, b4._lib.ptr(off), None))
    deg = np.diff(off).astype(np.int64)
    srcs = pick_sources(off, 1, args.sources)

    # ---- untimed pass: reached vertices / traversed edges per source ------
    e_trav, v_vis = [], []
    for s in srcs:
        depth, pred, st = bfs.bfs(graph, int(s), strategy=args.strategy,
                                  tuning=tuning_for(args.strategy, graph, info["arch"]))
        m = depth >= 0
        e_trav.append(int(deg[m].sum()))
        v_vis.append(int(m.sum()))
    log("per-source reached: mean V_vis=%.0f mean E_trav=%.0f" % (np.mean(v_vis), np.mean(e_trav)))

    head_tuning = tuning_for(args.strategy, graph, info["arch"])
    head = measure_strategy(bfs, b4, hip, graph, srcs, e_trav, v_vis, args.strategy, args.steps, args.warmup,
                            tuning=head_tuning)
    log("headline strategy %d: %s" % (args.strategy, json.dumps(head)))

    names = {0: "expand_contract", 1: "contract_expand", 2: "two_phase", 3: "hybrid", 4: "persistent",
             5: "direction_optimizing"}
    per_strategy = {names[args.strategy]: round(head["gteps"], 3)}
    if args.compare_steps > 0:
        # Each strategy in its default launch form (the reference's: fused for
        # expand-contract and contract-expand, test_bfs.cu:690-718), plus the
        # other form where both exist.
        for s, mode, name in ((0, 0, "expand_contract"), (0, 2, "expand_contract_iterative"),
                              (1, 0, "contract_expand"), (1, 2, "contract_expand_iterative"),
                              (2, 0, "two_phase"), (2, 1, "two_phase_persistent"), (3, 0, "hybrid"),
                              (5, 0, "direction_optimizing")):
            if s == args.strategy and mode == 0:
                continue
            try:
                r = measure_strategy(bfs, b4, hip, graph, srcs, e_trav, v_vis, s, args.compare_steps, 2, mode=mode,
                                     tuning=tuning_for(s, graph, info["arch"]))
            except b4.B40CError as ex:
                per_strategy[name] = None
                log("strategy %s: unavailable (%s)" % (name, ex))
                continue
            per_strategy[name] = round(r["gteps"], 3)
            log("strategy %s: %.3f GTEPS (%.3f ms/search)" % (name, r["gteps"], r["ms_per_search_enact"]))

    lattice = None
    if not args.no_lattice:
        try:
            lattice = bench_lattice(bfs, b4, hip, info)
        except b4.B40CError as ex:
            lattice = {"error": str(ex)}
            log("lattice workload failed:", ex)
    micro = None
    if not args.no_ingest:
        try:
            micro = bench_microbench(bfs, graph, srcs, args.strategy, head_tuning)
            log("microbench", json.dumps(micro))
        except b4.B40CError as ex:
            micro = {"error": str(ex)}
            log("microbench workload failed:", ex)
    prims = None
    if not args.no_ingest:
        try:
            prims = bench_primitives(hip)
            log("primitives", json.dumps(prims))
        except b4.B40CError as ex:
            prims = {"error": str(ex)}
            log("primitives workload failed:", ex)
    ingest = None
    if not args.no_ingest:
        try:
            ingest = bench_ingest()
            log("ingest", json.dumps(ingest))
        except b4.B40CError as ex:
            ingest = {"error": str(ex)}
            log("ingest workload failed:", ex)
    # The batched and served workloads run in child processes: they build
    # their own graph pools, and whatever happens there cannot take the
    # headline measurement down with it.
    batched = None if args.no_batched else run_child("batched")
    served = None if args.no_served else run_child("served")

    sort = None
    if not args.no_sort:
        try:
            sort = bench_sort(hip)
        except b4.B40CError as ex:
            sort = {"error": str(ex)}
            log("sort workload failed:", ex)

    # ---- CPU baseline: the serial oracle BFS, RMAT-16 designated ----------
    cpu = None
    if args.cpu_sources > 0:
        cpu = cpu_baseline_leg(b4, bfs, args.strategy, head_tuning, graph, off, srcs, args.scale,
                               sources22=args.cpu_sources)
        log("cpu baseline", cpu)

    achieved = head["achieved_gbs"]
    out_workload = "rmat%d_ef16_a0.57_b0.19_c0.19_sym_dedup_single_source_bfs" % args.scale
    out = {
        "metric": "BFS GTEPS (RMAT scale-22) + batched-query p50/p99 vs plain, 1 MI355X",
        "value": round(head["gteps"], 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": out_workload,
                   "vertices": n, "csr_edges": E, "sources": len(srcs), "strategy": names[args.strategy],
                   "tuning": head_tuning or "library defaults",
                   "step": "Reset + EnactSearch, depth + predecessor output"},
        "roofline": roofline_block(achieved, head["ms_per_search_enact"], out_workload, names[args.strategy]),
        "cpu_baseline": cpu,
        "lattice": lattice,
        "batched": batched,
        "served": served,
        "ingest": ingest,
        "primitives": prims,
        "sort": sort,
        "microbench": micro,
        "detail": {"gteps_enact_only": round(head["gteps_enact_only"], 3),
                   "ms_per_search_enact": round(head["ms_per_search_enact"], 4),
                   "b_alg_per_search": head["b_alg_per_search"],
                   "csr_build_s": round(build_s, 3), "per_strategy_gteps": per_strategy},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
