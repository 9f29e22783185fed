"""GPU parity on many small irregular graphs: random COO lists with
self-loops, duplicate edges, isolated vertices, directed and symmetric
forms, a single vertex and an edgeless graph, built through both the
reference-order COO path (duplicates kept, as CsrGraph::FromCoo keeps them)
and the deduplicating one.  Every enabled strategy runs from several
sources; depth must equal the serial oracle bit for bit and predecessors
must pass the reference's invariants (test_bfs.cu:387-438)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import back40computing_amd as b4  # noqa: E402

STRATEGIES = [b4.EXPAND_CONTRACT, b4.CONTRACT_EXPAND, b4.TWO_PHASE, b4.HYBRID, b4.DIRECTION_OPTIMIZING]


def random_coo(rng, n, m, symmetric):
    r = rng.integers(0, n, m).astype(np.int32)
    c = rng.integers(0, n, m).astype(np.int32)
    if m > 4:
        c[: m // 20] = r[: m // 20]                       # self-loops
        r[m // 20: m // 10] = r[m // 10: m // 10 + m // 20]  # duplicated rows
        c[m // 20: m // 10] = c[m // 10: m // 10 + m // 20]
    if n > 8:  # a block of isolated vertices
        keep = (r < n - n // 8) & (c < n - n // 8)
        r, c = r[keep], c[keep]
    if symmetric:
        r, c = np.concatenate([r, c]), np.concatenate([c, r])
    return r, c


CASES = [(seed, n, deg, sym, dedup) for seed, (n, deg) in enumerate(
    [(1, 0), (2, 1), (7, 3), (64, 2), (300, 1), (500, 6), (1000, 4), (3000, 8), (2048, 16), (5000, 2)])
    for sym in (False, True) for dedup in (False, True)]


@pytest.mark.parametrize("seed,n,deg,sym,dedup", CASES)
def test_random_graph_all_strategies(orc, seed, n, deg, sym, dedup):
    from back40computing_amd import bfs
    rng = np.random.default_rng(1000 + seed * 4 + 2 * sym + dedup)
    r, c = random_coo(rng, n, n * deg, sym)
    if dedup:
        g = bfs.Graph.from_coo_unique(n, r, c)
    else:
        g = bfs.Graph.from_coo(n, r, c)
    off, cols = g.to_host()
    srcs = sorted(set([0, n - 1, int(rng.integers(0, n))]))
    for s in srcs:
        ref, _, rsd = orc.bfs(off, cols, s, with_pred=False)
        for strat in STRATEGIES:
            print("random graph seed=%d n=%d sym=%d dedup=%d strategy=%d src=%d" % (seed, n, sym, dedup, strat, s),
                  flush=True)
            depth, pred, st = bfs.bfs(g, s, strategy=strat)
            np.testing.assert_array_equal(depth, ref)
            assert orc.check_pred(off, cols, ref, pred) == -1
            assert st["search_depth"] == rsd
