"""Runtime autotuner for the BFS enactors.

Replaces b40c's hand-set, compile-time per-architecture tuning policies
(KernelPolicy blocks such as b40c/graph/bfs/enactor_hybrid.cuh:569-655 and
b40c/radix_sort/autotuned_policy.cuh:58-659) with measurement: for a given
graph, strategy and sample of sources it runs a coordinate descent over the
runtime knobs of ``b40c_tuning`` (grid size, gather-tier thresholds,
direction-switch parameters, levels per host check), timing each candidate
with HIP events on the problem's stream.

Correctness of every accepted candidate: the default configuration's labels
for the sample sources are first proven to be BFS labels by
``verify_bfs_labels`` (local invariants that determine BFS depths uniquely,
checked over every edge of the graph), and a candidate is rejected unless
its depth labels equal those exactly and its predecessors pass the same
predecessor checks.

Results go to a JSON tuning table keyed by (gfx arch, strategy, graph class);
``lookup()`` returns the entry for a graph so callers can apply it with
``Enactor.set_tuning(**entry)``.
"""
import ctypes
import json
import math
import os
import time

import numpy as np

from . import _lib
from ._lib import check, lib

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_table.json")

# Candidate values per knob (coordinate descent visits them in this order).
SPACE = {
    "do_alpha": [1, 2, 3, 4, 6, 8, 14, 32, 64, 128, 256, 1024],
    "do_beta": [6, 12, 24, 48, 96, 192],
    "grid_gather_threshold": [1, 16, 32, 64, 128, 256, 512, 1024, 2048, 8192, 65536],
    "cta_gather_threshold": [128, 256, 512, 1024, 4096],
    "wave_gather_threshold": [16, 32, 64, 128],
    "grid_size": [64, 128, 256, 512, 1024, 1536, 2048, 4096],
    "levels_per_sync": [2, 4, 6, 7, 8, 12, 16, 32, 64],
    "saturation_quit": [64, 256, 512, 2048],
    "bottom_up_grid_size": [512, 1024, 1280, 1536, 2048, 4096],
    "persistent_wgs_per_cu": [1, 2, 3, 4, 6, 8],
}

# Knobs that matter per strategy.
KNOBS = {
    _lib.EXPAND_CONTRACT: ["persistent_wgs_per_cu", "grid_gather_threshold", "cta_gather_threshold",
                           "wave_gather_threshold", "grid_size", "levels_per_sync"],
    _lib.CONTRACT_EXPAND: ["persistent_wgs_per_cu", "grid_gather_threshold", "cta_gather_threshold",
                           "wave_gather_threshold", "grid_size", "levels_per_sync"],
    _lib.TWO_PHASE: ["persistent_wgs_per_cu", "grid_gather_threshold", "cta_gather_threshold",
                     "wave_gather_threshold", "grid_size", "levels_per_sync"],
    _lib.HYBRID: ["saturation_quit", "persistent_wgs_per_cu", "grid_gather_threshold", "cta_gather_threshold",
                  "wave_gather_threshold", "grid_size"],
    _lib.DIRECTION_OPTIMIZING: ["bottom_up_grid_size", "do_alpha", "do_beta", "grid_gather_threshold",
                                "grid_size", "levels_per_sync", "wave_gather_threshold"],
}


def verify_bfs_labels(offsets, cols, src, depth, pred, edge_keys=None, depth_verified=False):
    """True when (depth, pred) is a breadth-first labelling of the CSR graph
    from `src`, checked by invariants that determine BFS depths uniquely:
    depth[src] = 0 and pred[src] = -2; every other reached vertex v has
    pred[v] = u with an edge u -> v and depth[u] = depth[v] - 1; no edge
    leaves a reached vertex u for an unreached one or for one deeper than
    depth[u] + 1; unreached vertices have depth and pred -1.  (By induction
    on depth, the first two make depth[v] >= the true distance; the third
    makes it <= the distance and the reached set exactly the vertices
    reachable from src.)  `edge_keys` (row * n + col of every edge, sorted)
    may be passed to reuse it across sources.  With `depth_verified` the
    caller vouches that `depth` equals labels this function has already
    accepted, and only the predecessors are checked (no pass over the
    edges)."""
    off = np.asarray(offsets, np.int64)
    n = off.size - 1
    depth = np.asarray(depth)
    pred = np.asarray(pred)
    if depth.shape != (n,) or pred.shape != (n,) or not (0 <= src < n):
        return False
    if depth[src] != 0 or pred[src] != -2:
        return False
    reached = depth >= 0
    if np.any(depth[~reached] != -1) or np.any(pred[~reached] != -1):
        return False
    if not depth_verified:
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
        c = np.asarray(cols, np.int64)
        du = depth[rows]
        dv = depth[c]
        from_reached = du >= 0
        if np.any(dv[from_reached] < 0) or np.any(dv[from_reached] > du[from_reached] + 1):
            return False
    others = np.flatnonzero(reached)
    others = others[others != src]
    p = pred[others].astype(np.int64)
    if np.any((p < 0) | (p >= n)):
        return False
    if np.any(depth[p] != depth[others] - 1):
        return False
    if others.size == 0:
        return True
    if edge_keys is None:
        edge_keys = sorted_edge_keys(off, cols)
    if edge_keys.size == 0:
        return False
    want = p * n + others
    pos = np.minimum(np.searchsorted(edge_keys, want), edge_keys.size - 1)
    return bool(np.all(edge_keys[pos] == want))


def sorted_edge_keys(offsets, cols):
    """row * n + col of every edge, sorted (for verify_bfs_labels)."""
    off = np.asarray(offsets, np.int64)
    n = off.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    return np.sort(rows * n + np.asarray(cols, np.int64))


def graph_class(nodes, edges):
    """Coarse key: log2 vertex count and log2 average degree."""
    avg = edges / max(1, nodes)
    return "v%d_d%d" % (int(round(math.log2(max(2, nodes)))), int(round(math.log2(max(1.0, avg)))))


def _key(arch, strategy, nodes, edges):
    return "%s/%d/%s" % (arch.split(":")[0], strategy, graph_class(nodes, edges))


def load_table(path=TABLE_PATH):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(arch, strategy, nodes, edges, path=TABLE_PATH):
    """Tuning entry (dict of b40c_tuning fields) for this graph, or None."""
    ent = load_table(path).get(_key(arch, strategy, nodes, edges))
    return dict(ent["tuning"]) if ent else None


class _Timer:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.e0, self.e1 = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.L.hipEventCreate(ctypes.byref(self.e0)) == 0
        assert self.L.hipEventCreate(ctypes.byref(self.e1)) == 0

    def start(self, stream):
        assert self.L.hipEventRecord(self.e0, stream) == 0

    def stop_ms(self, stream):
        assert self.L.hipEventRecord(self.e1, stream) == 0
        assert self.L.hipEventSynchronize(self.e1) == 0
        ms = ctypes.c_float()
        assert self.L.hipEventElapsedTime(ctypes.byref(ms), self.e0, self.e1) == 0
        return ms.value


def measure(graph, strategy, sources, tuning, repeats=1, reference=None, pred_check=None):
    """Mean milliseconds per search (Reset + EnactSearch) for `tuning`; returns
    (ms, depths, preds).  Raises if the depth labels differ from `reference`
    or a predecessor array fails `pred_check(k, depth, pred)`."""
    from . import bfs
    prob = bfs.CsrProblem()
    check(prob.FromGraph(graph))
    en = bfs.ENACTORS[strategy]()
    en.set_tuning(**tuning)
    ft = en.GetFrontierType()
    stream = prob.stream()
    timer = _Timer()
    depths, preds = [], []
    # correctness pass (also warms up)
    for s in sources:
        check(prob.Reset(ft, 1.3))
        check(lib().b40c_enact_search(en.handle, prob.handle, int(s), 0, _lib.MODE_DEFAULT))
        d, p = prob.results()
        depths.append(d)
        preds.append(p)
    if reference is not None:
        for a, b in zip(depths, reference):
            if not np.array_equal(a, b):
                raise AssertionError("tuning %s changed the depth labels" % (tuning,))
    if pred_check is not None:
        for k, (d, p) in enumerate(zip(depths, preds)):
            if not pred_check(k, d, p):
                raise AssertionError("tuning %s produced an invalid predecessor" % (tuning,))
    total = 0.0
    for _ in range(repeats):
        timer.start(stream)
        for s in sources:
            check(prob.Reset(ft, 1.3))
            check(lib().b40c_enact_search(en.handle, prob.handle, int(s), 0, _lib.MODE_DEFAULT))
        total += timer.stop_ms(stream)
    prob.close()
    en.close()
    return total / (repeats * len(sources)), depths, preds


class LabelsInvalid(AssertionError):
    pass


def tune(graph, strategy, sources, rounds=2, repeats=2, log=None, knobs=None, space=None, start=None):
    """Coordinate descent from the library defaults (overridden by `start`,
    e.g. an earlier table entry).  Returns (tuning, ms, history).  The
    starting point's labels are verified as BFS labels first; every
    candidate must reproduce them."""
    from . import bfs
    space = space or SPACE
    en = bfs.ENACTORS[strategy]()
    best = {k: v for k, v in en.tuning.items() if k in space}
    en.close()
    if start:
        best.update({k: v for k, v in start.items() if k in space})
    off, cols = graph.to_host()
    keys = sorted_edge_keys(off, cols)
    best_ms, ref, ref_pred = measure(graph, strategy, sources, best, repeats)
    for s, d, p in zip(sources, ref, ref_pred):
        if not verify_bfs_labels(off, cols, int(s), d, p, keys):
            raise LabelsInvalid("default tuning gave labels that are not a BFS from %d" % int(s))

    def pred_ok(k, d, p):
        # d already equals the verified labels (measure compares it first).
        return verify_bfs_labels(off, cols, int(sources[k]), d, p, keys, depth_verified=True)

    history = [(dict(best), best_ms)]
    if log:
        log("default %s: %.4f ms/search" % (best, best_ms))
    for _ in range(rounds):
        improved = False
        for knob in (knobs or KNOBS[strategy]):
            for val in space[knob]:
                if val == best.get(knob):
                    continue
                cand = dict(best)
                cand[knob] = val
                try:
                    ms, _, _ = measure(graph, strategy, sources, cand, repeats, reference=ref, pred_check=pred_ok)
                except _lib.B40CError:
                    continue
                except AssertionError as ex:
                    # Wrong labels from a candidate are a library bug: never
                    # accept it, and say so.
                    history.append((dict(cand), None))
                    if log:
                        log("  REJECTED %s=%s: %s" % (knob, val, ex))
                    continue
                history.append((dict(cand), ms))
                better = ms < best_ms * 0.98
                if log:
                    log("  %s=%s -> %.4f ms/search%s" % (knob, val, ms, " (best)" if better else ""))
                if better:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    return best, best_ms, history


def tune_and_record(graph, strategy, sources, path=TABLE_PATH, log=None, **kw):
    from . import bfs
    arch = bfs.device_info()["arch"]
    t0 = time.time()
    best, ms, hist = tune(graph, strategy, sources, log=log, **kw)
    table = load_table(path)
    table[_key(arch, strategy, graph.nodes, graph.edges)] = {
        "tuning": best, "ms_per_search": round(ms, 5), "evaluated": len(hist),
        "nodes": graph.nodes, "edges": graph.edges, "sources": len(sources),
        "tune_seconds": round(time.time() - t0, 2),
    }
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return best, ms


# Radix-sort policies (b40c_sort_enactor_set_policy): keys per thread of a
# downsweep tile and workgroups per CU, the run-time form of the reference's
# per-architecture sort policies (radix_sort/autotuned_policy.cuh:190-230).
SORT_SPACE = [(items, wgs) for items in (8, 16) for wgs in (1, 2, 3, 4, 6, 8)]

SORT_KEY_DTYPES = {"u32": np.uint32, "i32": np.int32, "f32": np.float32, "u64": np.uint64, "i64": np.int64,
                   "f64": np.float64}


def _sort_key(arch, key_name, with_values, n):
    return "%s/sort/%s%s/n%d" % (arch.split(":")[0], key_name, "_pairs" if with_values else "",
                                 int(round(math.log2(max(2, n)))))


def lookup_sort(arch, key_name, with_values, n, path=TABLE_PATH):
    """(items_per_thread, wgs_per_cu) recorded for this sort class, or None."""
    ent = load_table(path).get(_sort_key(arch, key_name, with_values, n))
    return (ent["items_per_thread"], ent["wgs_per_cu"]) if ent else None


def _random_keys(rng, dt, n):
    if np.issubdtype(dt, np.floating):
        return rng.standard_normal(n).astype(dt)
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)


def tune_sort(n, key_name="u64", with_values=False, repeats=3, seed=5, log=None, space=None):
    """Times every policy of SORT_SPACE on n random keys (of the key type, with
    uint32 values if asked) with a sort enactor, best of `repeats` per policy.
    Every policy's output is compared with numpy's stable sort of the same
    input (keys, and the values carried with them) before its time counts.
    Returns (best policy, its ms, history)."""
    from . import radix_sort as rs
    dt = np.dtype(SORT_KEY_DTYPES[key_name])
    rng = np.random.default_rng(seed)
    keys = _random_keys(rng, dt, n)
    vals = np.arange(n, dtype=np.uint32) if with_values else None
    # Random normals hold no NaN and no signed zero, where numpy's order and
    # the bitwise key order could differ.
    order = np.argsort(keys, kind="stable")
    want_keys = keys[order]
    en = rs.Enactor()
    db = rs.DoubleBuffer(dt, n, with_values=with_values)
    timer = _Timer()
    best, best_ms, history = None, None, []
    try:
        for items, wgs in (space or SORT_SPACE):
            en.set_policy(items, wgs)
            ms = None
            for _ in range(repeats):
                db.selector = 0
                db.d_keys[0].upload(keys)
                if with_values:
                    db.d_values[0].upload(vals)
                t0 = time.perf_counter()
                check(en.Sort(db, n))
                dt_ms = (time.perf_counter() - t0) * 1e3
                ms = dt_ms if ms is None else min(ms, dt_ms)
            got = db.d_keys[db.selector].download(dt, n)
            ok = np.array_equal(got.view(np.uint8), want_keys.view(np.uint8))
            if with_values:
                ok = ok and np.array_equal(db.d_values[db.selector].download(np.uint32, n), order.astype(np.uint32))
            if not ok:
                history.append(((items, wgs), None))
                if log:
                    log("  REJECTED sort policy items=%d wgs_per_cu=%d: output differs from numpy" % (items, wgs))
                continue
            history.append(((items, wgs), ms))
            if log:
                log("  sort %s%s n=%d items=%d wgs_per_cu=%d: %.3f ms" % (key_name, "+v" if with_values else "", n,
                                                                        items, wgs, ms))
            if best_ms is None or ms < best_ms:
                best, best_ms = (items, wgs), ms
    finally:
        for a in db.d_keys + (db.d_values or []):
            a.free()
        en.close()
    return best, best_ms, history


def tune_sort_and_record(n, key_name="u64", with_values=False, path=TABLE_PATH, log=None, **kw):
    from . import bfs
    arch = bfs.device_info()["arch"]
    best, ms, hist = tune_sort(n, key_name, with_values, log=log, **kw)
    if best is None:
        raise AssertionError("no sort policy reproduced numpy's order")
    table = load_table(path)
    table[_sort_key(arch, key_name, with_values, n)] = {
        "items_per_thread": best[0], "wgs_per_cu": best[1], "ms": round(ms, 4), "n": n,
        "evaluated": len(hist), "all_ms": {"%d/%d" % p: (round(t, 4) if t else None) for p, t in hist},
    }
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return best, ms


def main(argv=None):
    """Tune strategies on a synthetic RMAT or lattice graph and write the table.

    python -m back40computing_amd.autotune --graph rmat --scale 22 --strategies 5,0 --out table.json
    """
    import argparse
    import sys
    from . import bfs
    from .sources import pick_sources
    ap = argparse.ArgumentParser()
    ap.add_argument("--graph", choices=["rmat", "lattice"], default="rmat")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--width", type=int, default=200)
    ap.add_argument("--strategies", default="5")
    ap.add_argument("--sources", type=int, default=8)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--knobs", default=None, help="comma-separated knobs to search (default: the strategy's set)")
    ap.add_argument("--from-table", action="store_true",
                    help="start from the entry already recorded for the graph class, if any")
    ap.add_argument("--sort", default=None,
                    help="tune the radix sort instead: KEY[,pairs] (KEY in u32 i32 f32 u64 i64 f64), e.g. u64")
    ap.add_argument("--sort-n", type=int, default=1 << 27, help="keys per sort when tuning the radix sort")
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.sort:
        parts = args.sort.split(",")
        best, ms = tune_sort_and_record(args.sort_n, parts[0], "pairs" in parts[1:], path=args.out, log=log)
        log("sort %s n=%d best items=%d wgs_per_cu=%d: %.3f ms" % (args.sort, args.sort_n, best[0], best[1], ms))
        return 0

    if args.graph == "rmat":
        g = bfs.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, 0)
    else:
        g = bfs.Graph.lattice3d(args.width, 0)
    off = np.empty(g.nodes + 1, np.int32)
    check(lib().b40c_graph_copy_to_host(g.handle, _lib.ptr(off), None))
    srcs = pick_sources(off, 7, args.sources)
    knobs = args.knobs.split(",") if args.knobs else None
    arch = bfs.device_info()["arch"]
    for s in (int(x) for x in args.strategies.split(",")):
        log("tuning strategy %d on %s (V=%d E=%d)" % (s, args.graph, g.nodes, g.edges))
        start = lookup(arch, s, g.nodes, g.edges) if args.from_table else None
        best, ms = tune_and_record(g, s, srcs, path=args.out, log=log, rounds=args.rounds, knobs=knobs,
                                   start=start)
        log("strategy %d best %.4f ms/search: %s" % (s, ms, best))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
