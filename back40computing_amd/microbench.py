"""Per-stage BFS microbenchmarks on the GPU (libb40c_mi355x.so).

A level-synchronous BFS split into its status lookup and neighbour gather
stages; each stage runs once as a bench pass (culls and reads only, on private
copies of the visited state, counting the survivors of every cull) and once for
real, and every launch is timed with HIP events.  This is what the reference's
microbenchmark drivers measure: ``microbench_bfs.cu`` with
``EnactorGatherLookup`` (test/bfs/microbench/enactor_gather_lookup.cuh:
152-382, mode ``"gather_lookup"``) and ``microbench_cull.cu`` with
``EnactorCull`` (test/bfs/microbench/enactor_cull.cuh:152-385, mode
``"cull"``).  The per-level table printed by ``format_table`` carries the
reference drivers' columns (iteration, culled-queue and expansion-queue
lengths) plus the times.

    python -m back40computing_amd.microbench --rmat 22 --mode cull
"""
import argparse
import ctypes
import json
import sys

import numpy as np

from . import _lib
from ._lib import MbLevel, check, lib, ptr

MODES = {"gather_lookup": _lib.MB_GATHER_LOOKUP, "cull": _lib.MB_CULL}

# Algorithmic bytes per unit of each stage (int32 ids, int32 CSR offsets):
#   lookup: 4 B id read + 4 B visited-mask word probe per input id, and per
#           discovered vertex 4 B label store + 8 B row offsets + 8 B pair store;
#   offsets (the gather's set-up): 4 B row-length read in each of the scan's
#           two passes + 4 B offset store per input vertex;
#   gather: 8 B (row start, output offset) read per input vertex, 4 B column
#           read + 4 B id store per gathered neighbour (the bench pass stores
#           nothing).
LOOKUP_B_PER_ID = 8
LOOKUP_B_PER_VERTEX = 20
OFFSETS_B_PER_VERTEX = 12
GATHER_B_PER_VERTEX = 8
GATHER_B_PER_EDGE = 8
BENCH_GATHER_B_PER_EDGE = 4


def search(graph, src, mode="gather_lookup", want_depth=True, max_levels=4096):
    """Runs the microbenchmark search from ``src``; returns (depth or None,
    list of per-iteration dicts of b40c_mb_level fields)."""
    if mode not in MODES:
        raise ValueError("mode must be one of %s" % sorted(MODES))
    depth = np.empty(graph.nodes, np.int32) if want_depth else None
    while True:
        recs = (MbLevel * max_levels)()
        cnt = ctypes.c_int32()
        check(lib().b40c_microbench_search(graph.handle, int(src), MODES[mode],
                                           ptr(depth) if depth is not None else None, recs, max_levels,
                                           ctypes.byref(cnt)))
        if cnt.value <= max_levels:
            return depth, [recs[i].as_dict() for i in range(cnt.value)]
        max_levels = cnt.value


def summarize(levels):
    """Totals per stage: milliseconds, algorithmic bytes, GB/s and rates."""
    out = {}
    pairs_in = sum(r["discovered"] for r in levels if r["gathered"])
    lookup_in = sum(r["lookup_in"] for r in levels)
    found = sum(r["discovered"] for r in levels if r["iteration"] > 0)
    gathered = sum(r["gathered"] for r in levels)
    stages = {
        "lookup": (sum(r["lookup_ms"] for r in levels), LOOKUP_B_PER_ID * lookup_in + LOOKUP_B_PER_VERTEX * found),
        "bench_lookup": (sum(r["bench_lookup_ms"] for r in levels), LOOKUP_B_PER_ID * lookup_in),
        "offsets": (sum(r["offsets_ms"] for r in levels), OFFSETS_B_PER_VERTEX * pairs_in),
        "gather": (sum(r["gather_ms"] for r in levels), GATHER_B_PER_VERTEX * pairs_in + GATHER_B_PER_EDGE * gathered),
        "bench_gather": (sum(r["bench_gather_ms"] for r in levels),
                         GATHER_B_PER_VERTEX * pairs_in + BENCH_GATHER_B_PER_EDGE * gathered),
    }
    for name, (ms, nbytes) in stages.items():
        out[name] = {"ms": round(ms, 4), "bytes": int(nbytes),
                     "gb_s": round(nbytes / (ms * 1e6), 1) if ms > 0 else None}
    lk = out["lookup"]["ms"]
    ga = out["gather"]["ms"]
    out["lookup"]["gids_s"] = round(lookup_in / (lk * 1e6), 2) if lk > 0 else None
    out["gather"]["gedges_s"] = round(gathered / (ga * 1e6), 2) if ga > 0 else None
    out["ids_looked_up"] = int(lookup_in)
    out["edges_gathered"] = int(gathered)
    out["vertices_found"] = int(found) + 1
    bm = [r for r in levels if r["bitmask_pass"] >= 0]
    if lookup_in:
        out["bitmask_pass_frac"] = round(sum(r["bitmask_pass"] for r in bm) / lookup_in, 4)
        out["label_pass_frac"] = round(sum(r["label_pass"] for r in bm) / lookup_in, 4)
        loc = [r for r in bm if r["local_pass"] >= 0]
        if loc:
            out["local_pass_frac"] = round(sum(r["local_pass"] for r in loc) / lookup_in, 4)
    return out


def format_table(levels, mode="gather_lookup"):
    """CSV per iteration with the reference drivers' queue columns and the times."""
    cols = ["iteration", "lookup_in", "bitmask_pass", "label_pass"]
    if mode == "cull":
        cols.append("local_pass")
    cols += ["discovered", "gathered", "bench_lookup_ms", "lookup_ms", "offsets_ms", "bench_gather_ms", "gather_ms"]
    lines = [", ".join(cols)]
    for r in levels:
        lines.append(", ".join(("%.4f" % r[c]) if c.endswith("_ms") else str(r[c]) for c in cols))
    return "\n".join(lines)


def main(argv=None):
    from . import bfs, graph_io
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    g = ap.add_mutually_exclusive_group()
    g.add_argument("--rmat", type=int, help="RMAT scale (edge factor --edge-factor)")
    g.add_argument("--lattice", type=int, help="3-D lattice width")
    g.add_argument("--file", help="graph file (--format market|metis|dimacs)")
    ap.add_argument("--format", default="market", choices=sorted(graph_io.FORMATS))
    ap.add_argument("--undirected", action="store_true")
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--src", type=int, default=0)
    ap.add_argument("--mode", default="gather_lookup", choices=sorted(MODES))
    ap.add_argument("--repeat", type=int, default=2,
                    help="searches to run; the last one is reported (the first pays one-time kernel loads)")
    ap.add_argument("--json", action="store_true", help="print the summary as one JSON line")
    a = ap.parse_args(argv)
    if a.file:
        graph = graph_io.read_graph(a.file, a.format, undirected=a.undirected)
    elif a.lattice:
        graph = bfs.Graph.lattice3d(a.lattice)
    else:
        graph = bfs.Graph.rmat(a.rmat or 20, a.edge_factor, 0.57, 0.19, 0.19, 0)
    for _ in range(max(1, a.repeat)):
        _, levels = search(graph, a.src, a.mode, want_depth=False)
    summary = summarize(levels)
    if a.json:
        print(json.dumps(summary))
    else:
        print(format_table(levels, a.mode))
        print(json.dumps(summary, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
