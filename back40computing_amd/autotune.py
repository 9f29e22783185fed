"""Runtime autotuner for the BFS enactors.

Replaces b40c's hand-set, compile-time per-architecture tuning policies
(KernelPolicy blocks such as b40c/graph/bfs/enactor_hybrid.cuh:569-655 and
b40c/radix_sort/autotuned_policy.cuh:58-659) with measurement: for a given
graph, strategy and sample of sources it runs a coordinate descent over the
runtime knobs of ``b40c_tuning`` (grid size, gather-tier thresholds,
direction-switch parameters, levels per host check), timing each candidate
with HIP events on the problem's stream and rejecting any candidate whose
depth labels differ from the default configuration's.

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
    "grid_size": [256, 512, 1024, 1536, 2048, 4096],
    "levels_per_sync": [2, 4, 6, 7, 8, 12, 16, 32, 64],
    "saturation_quit": [64, 256, 512, 2048],
    "bottom_up_grid_size": [512, 1024, 1280, 1536, 2048, 4096],
}

# Knobs that matter per strategy.
KNOBS = {
    _lib.EXPAND_CONTRACT: ["grid_gather_threshold", "cta_gather_threshold", "wave_gather_threshold",
                           "grid_size", "levels_per_sync"],
    _lib.CONTRACT_EXPAND: ["grid_gather_threshold", "cta_gather_threshold", "wave_gather_threshold",
                           "grid_size", "levels_per_sync"],
    _lib.TWO_PHASE: ["grid_gather_threshold", "cta_gather_threshold", "wave_gather_threshold",
                     "grid_size", "levels_per_sync"],
    _lib.HYBRID: ["saturation_quit", "grid_gather_threshold", "cta_gather_threshold", "wave_gather_threshold",
                  "grid_size"],
    _lib.DIRECTION_OPTIMIZING: ["bottom_up_grid_size", "do_alpha", "do_beta", "grid_gather_threshold",
                                "grid_size", "levels_per_sync", "wave_gather_threshold"],
}


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


def measure(graph, strategy, sources, tuning, repeats=1, reference=None):
    """Mean milliseconds per search (Reset + EnactSearch) for `tuning`; returns
    (ms, depths).  Raises if the labels differ from `reference`."""
    from . import bfs
    prob = bfs.CsrProblem()
    check(prob.FromGraph(graph))
    en = bfs.ENACTORS[strategy]()
    en.set_tuning(**tuning)
    ft = en.GetFrontierType()
    stream = prob.stream()
    timer = _Timer()
    depths = []
    # correctness pass (also warms up)
    for s in sources:
        check(prob.Reset(ft, 1.3))
        check(lib().b40c_enact_search(en.handle, prob.handle, int(s), 0, _lib.MODE_DEFAULT))
        d, _ = prob.results()
        depths.append(d)
    if reference is not None:
        for a, b in zip(depths, reference):
            if not np.array_equal(a, b):
                raise AssertionError("tuning %s changed the depth labels" % (tuning,))
    total = 0.0
    for _ in range(repeats):
        timer.start(stream)
        for s in sources:
            check(prob.Reset(ft, 1.3))
            check(lib().b40c_enact_search(en.handle, prob.handle, int(s), 0, _lib.MODE_DEFAULT))
        total += timer.stop_ms(stream)
    prob.close()
    en.close()
    return total / (repeats * len(sources)), depths


def tune(graph, strategy, sources, rounds=2, repeats=2, log=None):
    """Coordinate descent from the library defaults.  Returns (tuning, ms, history)."""
    from . import bfs
    en = bfs.ENACTORS[strategy]()
    best = {k: v for k, v in en.tuning.items() if k in SPACE}
    en.close()
    best_ms, ref = measure(graph, strategy, sources, best, repeats)
    history = [(dict(best), best_ms)]
    if log:
        log("default %s: %.4f ms/search" % (best, best_ms))
    for _ in range(rounds):
        improved = False
        for knob in KNOBS[strategy]:
            for val in SPACE[knob]:
                if val == best.get(knob):
                    continue
                cand = dict(best)
                cand[knob] = val
                try:
                    ms, _ = measure(graph, strategy, sources, cand, repeats, reference=ref)
                except _lib.B40CError:
                    continue
                history.append((dict(cand), ms))
                if ms < best_ms * 0.98:
                    best, best_ms, improved = cand, ms, True
                    if log:
                        log("  %s=%s -> %.4f ms/search" % (knob, val, ms))
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
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.graph == "rmat":
        g = bfs.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, 0)
    else:
        g = bfs.Graph.lattice3d(args.width, 0)
    off = np.empty(g.nodes + 1, np.int32)
    check(lib().b40c_graph_copy_to_host(g.handle, _lib.ptr(off), None))
    srcs = pick_sources(off, 7, args.sources)
    for s in (int(x) for x in args.strategies.split(",")):
        log("tuning strategy %d on %s (V=%d E=%d)" % (s, args.graph, g.nodes, g.edges))
        best, ms = tune_and_record(g, s, srcs, path=args.out, log=log, rounds=args.rounds)
        log("strategy %d best %.4f ms/search: %s" % (s, ms, best))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
