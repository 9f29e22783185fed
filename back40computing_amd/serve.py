"""Batched and served BFS workloads, and the request loop that serves them.

* ``build_pool`` — RMAT graphs built on the GPU and kept resident.
* ``ContinuousServer`` — the serving loop: requests that arrive while others
  are on the GPU are admitted at the next level boundary as graph slots and
  label memory free up (``MultiSourceBFS``), so their work overlaps on the
  device; latency is arrival -> labels ready on the device.
* ``PlainServer`` — the comparator: the same requests served one at a time,
  in arrival order, by the single-source enactor.

Both servers replay a precomputed arrival trace against the wall clock, so
queueing delay is real.
"""
import time

import numpy as np

from . import _lib
from .batch import MultiGraph, MultiSourceBFS
from .bfs import CsrProblem, ENACTORS, Graph
from .sources import pick_sources, splitmix64


def _uniform(seed, k):
    return (splitmix64((seed * 0x9E3779B97F4A7C15 + k) & ((1 << 64) - 1)) >> 11) / float(1 << 53)


def build_pool(scales, edge_factor=16, a=0.57, b=0.19, c=0.19, seed0=1000):
    """One RMAT graph per scale (seeded seed0 + index); returns graphs and host offsets."""
    graphs, offsets = [], []
    for i, s in enumerate(scales):
        g = Graph.rmat(int(s), edge_factor, a, b, c, seed0 + i)
        off = np.empty(g.nodes + 1, np.int32)
        _lib.check(_lib.lib().b40c_graph_copy_to_host(g.handle, _lib.ptr(off), None))
        graphs.append(g)
        offsets.append(off)
    return graphs, offsets


def batched_workload(num_graphs=256, per_graph=4, scale_lo=10, scale_hi=18, seed=4):
    """Graph scales integer-uniform in [scale_lo, scale_hi], `per_graph` seeded
    sources of nonzero degree each."""
    scales = [scale_lo + int(_uniform(seed, k) * (scale_hi - scale_lo + 1)) for k in range(num_graphs)]
    graphs, offsets = build_pool(scales, seed0=seed * 100000)
    queries = []
    for g, off in enumerate(offsets):
        for s in pick_sources(off, seed * 7919 + g, per_graph):
            queries.append((g, int(s)))
    return graphs, offsets, scales, queries


def poisson_trace(rate_qps, duration_s, choose, seed=9):
    """Arrival times (s) of a Poisson process and the query chosen for each."""
    t, k, out = 0.0, 0, []
    while True:
        u = max(_uniform(seed, 2 * k), 1e-12)
        t += -np.log(u) / rate_qps
        if t >= duration_s:
            return out
        out.append((t, choose(_uniform(seed, 2 * k + 1), k)))
        k += 1


class ContinuousServer:
    def __init__(self, multigraph, label_budget_bytes, max_queries=1 << 20, tuning=None, log=None, stall_s=10.0,
                 levels_per_step=0):
        self.eng = MultiSourceBFS(multigraph, label_budget_bytes, max_queries)
        if tuning:
            self.eng.set_tuning(**tuning)
        self.eng.set_levels_per_step(levels_per_step)
        self.mg = multigraph
        self.log = log
        self.stall_s = stall_s

    def serve(self, trace):
        """trace: [(arrival_s, (graph, src))] sorted.  Returns (latencies_s, wall_s);
        latency = arrival -> the query's labels are complete in device memory."""
        t0 = time.perf_counter()
        pending = list(trace)
        idx = 0
        qid_of = {}
        lat = {}
        in_flight = 0
        last_progress = t0
        last_log = t0
        steps = 0
        while idx < len(pending) or in_flight > 0:
            now = time.perf_counter() - t0
            while idx < len(pending) and pending[idx][0] <= now:
                g, s = pending[idx][1]
                qid_of[self.eng.submit(g, s)] = idx
                idx += 1
                in_flight += 1
            if in_flight == 0:
                time.sleep(max(0.0, min(pending[idx][0] - now, 0.001)))
                continue
            running, waiting = self.eng.step()
            steps += 1
            wall_now = time.perf_counter()
            done_t = wall_now - t0
            finished = self.eng.poll()
            for q in finished:
                i = qid_of[q]
                lat[i] = done_t - pending[i][0]
                in_flight -= 1
                self.eng.release(q)  # labels were complete on the device; free the slot
            if finished or running:
                last_progress = wall_now
            elif wall_now - last_progress > self.stall_s:
                raise RuntimeError("serving loop made no progress for %.1f s (%d waiting, %d running)"
                                   % (self.stall_s, waiting, running))
            if self.log and wall_now - last_log > 0.5:
                self.log("  serve t=%.2fs submitted=%d done=%d running=%d waiting=%d steps=%d" % (
                    done_t, idx, len(lat), running, waiting, steps))
                last_log = wall_now
        wall = time.perf_counter() - t0
        return np.array([lat[i] for i in range(len(pending))]), wall


class PlainServer:
    """FIFO server over the single-source enactor: one resident problem and
    one enactor per graph (created on first use), so each graph keeps its
    tuned knobs and its captured level loop."""

    def __init__(self, graphs, strategy=_lib.DIRECTION_OPTIMIZING, tuning_for=None):
        self.graphs = graphs
        self.strategy = strategy
        self.tuning_for = tuning_for
        self.problems = {}
        self.enactors = {}

    def _slot(self, g):
        p = self.problems.get(g)
        if p is None:
            p = CsrProblem()
            _lib.check(p.FromGraph(self.graphs[g]))
            en = ENACTORS[self.strategy]()
            t = self.tuning_for(g) if self.tuning_for else None
            if t:
                en.set_tuning(**t)
            self.problems[g] = p
            self.enactors[g] = en
        return p, self.enactors[g]

    def run_one(self, g, src, want=False):
        p, en = self._slot(g)
        _lib.check(p.Reset(en.GetFrontierType(), 1.3))
        _lib.check(_lib.lib().b40c_enact_search(en.handle, p.handle, int(src), 0, _lib.MODE_DEFAULT))
        if want:
            return p.results()
        return None

    def serve(self, trace, log=None):
        t0 = time.perf_counter()
        last_log = t0
        lat = []
        for arr, (g, s) in trace:
            now = time.perf_counter() - t0
            if now < arr:
                time.sleep(arr - now)
            self.run_one(g, s)  # returns once the labels are complete on the device
            lat.append(time.perf_counter() - t0 - arr)
            if log and time.perf_counter() - last_log > 0.5:
                log("  plain t=%.2fs done=%d/%d" % (time.perf_counter() - t0, len(lat), len(trace)))
                last_log = time.perf_counter()
        return np.array(lat), time.perf_counter() - t0

    def close(self):
        for p in self.problems.values():
            p.close()
        for en in self.enactors.values():
            en.close()
