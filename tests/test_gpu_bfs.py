"""GPU parity: every strategy and launch mode against the serial oracle.

Depth labels are compared bit-exactly; predecessors are checked with the
reference's invariants (test/bfs/test_bfs.cu:387-438) plus source = -2;
search depth must equal the oracle's max label + 1."""
import numpy as np
import pytest

from helpers import csr_from_edges, hand_graphs

pytestmark = pytest.mark.gpu

import back40computing_amd as b4  # noqa: E402

ALL_MODES = [
    (b4.EXPAND_CONTRACT, b4.MODE_ITERATIVE),
    (b4.EXPAND_CONTRACT, b4.MODE_FUSED),
    (b4.CONTRACT_EXPAND, b4.MODE_ITERATIVE),
    (b4.CONTRACT_EXPAND, b4.MODE_FUSED),
    (b4.TWO_PHASE, b4.MODE_ITERATIVE),
    (b4.TWO_PHASE, b4.MODE_FUSED),
    (b4.HYBRID, b4.MODE_DEFAULT),
    (b4.PERSISTENT, b4.MODE_DEFAULT),
    (b4.DIRECTION_OPTIMIZING, b4.MODE_DEFAULT),
]


def params():
    return [pytest.param(sm, id="%d-%d" % sm) for sm in ALL_MODES]


STRATEGY_MODES = params()


@pytest.fixture(scope="module")
def bfsmod():
    from back40computing_amd import bfs
    return bfs


def run(bfsmod, graph, src, strategy, mode, tuning=None, queue_sizing=1.3):
    print("bfs strategy=%d mode=%d src=%d n=%d" % (strategy, mode, src, graph.nodes), flush=True)
    return bfsmod.bfs(graph, src, strategy=strategy, mode=mode, tuning=tuning, queue_sizing=queue_sizing)


def check(orc, off, cols, src, depth, pred, stats):
    ref, _, sd = orc.bfs(off, cols, int(src), with_pred=False)
    np.testing.assert_array_equal(depth, ref)
    assert orc.check_pred(off, cols, ref, pred) == -1
    assert stats["search_depth"] == sd


@pytest.fixture(scope="module")
def rmat16(bfsmod, orc):
    g = bfsmod.Graph.rmat(16, 16, 0.57, 0.19, 0.19, 0)
    off, cols = g.to_host()
    srcs = orc.pick_sources(off, 1, 3)
    return g, off, cols, srcs


@pytest.mark.parametrize("g", hand_graphs(), ids=lambda g: g["name"])
@pytest.mark.parametrize("sm", STRATEGY_MODES)
def test_hand_graphs(bfsmod, orc, g, sm):
    off, cols = csr_from_edges(g["n"], g["edges"], g["undirected"])
    graph = bfsmod.Graph.from_host_csr(off, cols)
    for case in g["cases"]:
        depth, pred, st = run(bfsmod, graph, case["src"], *sm)
        assert depth.tolist() == case["depth"]
        assert st["search_depth"] == case["search_depth"]
        assert orc.check_pred(off, cols, depth, pred) == -1


@pytest.mark.parametrize("sm", STRATEGY_MODES)
def test_rmat16_all_strategies(bfsmod, orc, rmat16, sm):
    g, off, cols, srcs = rmat16
    for s in srcs:
        depth, pred, st = run(bfsmod, g, int(s), *sm)
        check(orc, off, cols, s, depth, pred, st)


@pytest.mark.parametrize("sm", STRATEGY_MODES)
def test_lattice_closed_form(bfsmod, orc, sm):
    w = 24
    g, perm = bfsmod.Graph.lattice3d(w, 5, want_perm=True)
    off, cols = g.to_host()
    for s in (0, w ** 3 // 2 + 7):
        depth, pred, st = run(bfsmod, g, s, *sm)
        np.testing.assert_array_equal(depth, orc.lattice_depth(w, perm, s))
        assert orc.check_pred(off, cols, depth, pred) == -1


@pytest.mark.parametrize("sm", STRATEGY_MODES)
def test_gather_tiers(bfsmod, orc, sm):
    # A star whose hub exceeds the workgroup tier, mid-degree vertices for the
    # wave tier and a long tail for the fine tier; thresholds lowered so each
    # tier is exercised on a small graph.
    rng = np.random.default_rng(3)
    n = 20000
    edges = [(0, i) for i in range(1, 9000)]
    for h in range(9000, 9100):
        edges += [(h, int(x)) for x in rng.integers(9100, n, 100)]
    edges += [(int(a), int(b)) for a, b in rng.integers(1, n, (30000, 2))]
    off, cols = csr_from_edges(n, edges, undirected=True)
    g = bfsmod.Graph.from_host_csr(off, cols)
    for tuning in (None, {"cta_gather_threshold": 300, "wave_gather_threshold": 8},
                   {"grid_gather_threshold": 50, "cta_gather_threshold": 20, "wave_gather_threshold": 4},
                   {"grid_gather_threshold": 1 << 30, "cta_gather_threshold": 1 << 30}):
        for s in (0, 9050, 15000):
            depth, pred, st = run(bfsmod, g, s, *sm, tuning=tuning)
            check(orc, off, cols, s, depth, pred, st)


def test_direction_optimizing_hub_past_grid_tier_levels(bfsmod, orc):
    """Symmetric graphs get the grid tier only on the first levels of each
    captured block; a hub reached top-down later must be gathered inside the
    level kernel.  A 6-edge path leads from vertex 0 to a hub of 9,000
    leaves, behind which lies a random region large enough (with alpha 1)
    to keep the search top-down when the hub is expanded at level 6."""
    rng = np.random.default_rng(11)
    n = 40000
    edges = [(i, i + 1) for i in range(6)]
    edges += [(6, i) for i in range(7, 9007)]
    edges += [(9006, 9007)]
    edges += [(int(a), int(b)) for a, b in rng.integers(9007, n, (60000, 2))]
    off, cols = csr_from_edges(n, edges, undirected=True)
    g = bfsmod.Graph.from_host_csr(off, cols)
    for tuning in ({"do_alpha": 1},
                   {"do_alpha": 1, "grid_gather_threshold": 50, "cta_gather_threshold": 20,
                    "wave_gather_threshold": 4}):
        for s in (0, 3, 9006):
            depth, pred, st = run(bfsmod, g, s, b4.DIRECTION_OPTIMIZING, b4.MODE_DEFAULT, tuning=tuning)
            check(orc, off, cols, s, depth, pred, st)
        assert int(depth[6]) >= 0


@pytest.mark.parametrize("sm", STRATEGY_MODES)
def test_grid_tier_many_chunks_per_workgroup(bfsmod, orc, rmat16, sm):
    """Grid tier with a low threshold on 2 or 3 workgroups: each workgroup
    owns thousands of 256-edge chunks, so the pair tier reserves several
    batches per workgroup and the last round of in-flight chunks is partial."""
    g, off, cols, srcs = rmat16
    for grid in (2, 3):
        tuning = {"grid_gather_threshold": 16, "grid_size": grid, "bottom_up_grid_size": grid}
        for s in srcs[:2]:
            depth, pred, st = run(bfsmod, g, int(s), *sm, tuning=tuning)
            check(orc, off, cols, int(s), depth, pred, st)


@pytest.mark.parametrize("sm", STRATEGY_MODES)
def test_reuse_problem_across_sources(bfsmod, orc, rmat16, sm):
    g, off, cols, srcs = rmat16
    strategy, mode = sm
    prob = bfsmod.CsrProblem()
    assert prob.FromGraph(g) == 0
    en = bfsmod.ENACTORS[strategy]()
    for s in srcs:
        assert prob.Reset(en.GetFrontierType(), 1.3) == 0
        assert b4.lib().b40c_enact_search(en.handle, prob.handle, int(s), 0, mode) == 0
        depth, pred = prob.results()
        q, sd, duty = en.GetStatistics()
        check(orc, off, cols, s, depth, pred, {"search_depth": sd})
        assert q > 0


def test_enact_requires_reset(bfsmod, rmat16):
    g, off, cols, srcs = rmat16
    prob = bfsmod.CsrProblem()
    assert prob.FromGraph(g) == 0
    en = bfsmod.EnactorHybrid()
    assert en.EnactSearch(prob, int(srcs[0])) == b4._lib.ERR_INVALID_ARG
    assert prob.Reset(en.GetFrontierType()) == 0
    assert en.EnactSearch(prob, -1) == b4._lib.ERR_INVALID_ARG
    assert en.EnactSearch(prob, int(srcs[0])) == 0
    # a second search without Reset is refused
    assert en.EnactSearch(prob, int(srcs[0])) == b4._lib.ERR_INVALID_ARG
    # a strategy needing edge frontiers refuses a vertex-frontier reset
    assert prob.Reset(b4.VERTEX_FRONTIERS) == 0
    assert bfsmod.EnactorContractExpand().EnactSearch(prob, int(srcs[0])) == b4._lib.ERR_INVALID_ARG


@pytest.mark.parametrize("strategy,mode", [(b4.CONTRACT_EXPAND, b4.MODE_ITERATIVE),
                                           (b4.TWO_PHASE, b4.MODE_ITERATIVE),
                                           (b4.HYBRID, b4.MODE_DEFAULT),
                                           (b4.EXPAND_CONTRACT, b4.MODE_ITERATIVE),
                                           (b4.CONTRACT_EXPAND, b4.MODE_FUSED),
                                           (b4.TWO_PHASE, b4.MODE_FUSED),
                                           (b4.EXPAND_CONTRACT, b4.MODE_FUSED)])
def test_forced_queue_overflow(bfsmod, rmat16, strategy, mode):
    """A tiny queue-sizing factor must surface as QUEUE_OVERFLOW, never as
    silent corruption (enactor_contract_expand.cuh:319-324)."""
    g, off, cols, srcs = rmat16
    with pytest.raises(b4.B40CError) as ei:
        run(bfsmod, g, int(srcs[0]), strategy, mode, queue_sizing=0.001)
    assert ei.value.status == b4._lib.ERR_QUEUE_OVERFLOW


def test_statistics_and_instrumentation(bfsmod, orc, rmat16):
    g, off, cols, srcs = rmat16
    for cls in (bfsmod.EnactorExpandContract, bfsmod.EnactorTwoPhase, bfsmod.EnactorContractExpand):
        prob = bfsmod.CsrProblem()
        assert prob.FromGraph(g) == 0
        en = cls(instrument=True)
        assert prob.Reset(en.GetFrontierType()) == 0
        assert en.EnactSearch(prob, int(srcs[0])) == 0
        depth, pred = prob.results()
        q, sd, duty = en.GetStatistics()
        ref, _, rsd = orc.bfs(off, cols, int(srcs[0]), with_pred=False)
        np.testing.assert_array_equal(depth, ref)
        assert sd == rsd
        assert 0.0 < duty <= 1.0 + 1e-9
        assert q >= int((ref >= 0).sum())


def test_fused_entry_points(bfsmod, orc, rmat16):
    """EnactFusedSearch of each enactor class and the persistent enactor's
    EnactSearch: one cooperative launch each, labels equal to the oracle's,
    and the problem reused for a second source (the grid barrier's epochs
    continue from the first search)."""
    g, off, cols, srcs = rmat16
    for cls in (bfsmod.EnactorExpandContract, bfsmod.EnactorContractExpand, bfsmod.EnactorTwoPhase,
                bfsmod.EnactorPersistent):
        prob = bfsmod.CsrProblem()
        assert prob.FromGraph(g) == 0
        en = cls()
        for s in srcs[:2]:
            print("fused entry %s src=%d" % (cls.__name__, int(s)), flush=True)
            assert prob.Reset(en.GetFrontierType()) == 0
            if cls is bfsmod.EnactorPersistent:
                assert en.EnactSearch(prob, int(s)) == 0
            else:
                assert en.EnactFusedSearch(prob, int(s)) == 0
            depth, pred = prob.results()
            q, sd, duty = en.GetStatistics()
            check(orc, off, cols, s, depth, pred, {"search_depth": sd})


@pytest.mark.parametrize("light", [True, False], ids=["light", "fenced"])
def test_gbarrier_zz_visibility(bfsmod, light):
    """Words written before the software grid barrier are seen by other
    workgroups after it, in both barrier modes (light: the words go through
    agent-scope accesses; fenced: plain accesses ordered by the barrier's
    release/acquire).  Indices are fixed and in bounds; values are only
    compared."""
    info = bfsmod.device_info()
    for grid in (info["num_cus"] // 4, info["num_cus"], 2 * info["num_cus"], 6 * info["num_cus"]):
        print("gbarrier_check grid=%d light=%s" % (grid, light), flush=True)
        assert bfsmod.gbarrier_check(grid, 200, light) == 0


@pytest.mark.parametrize("light", [True, False], ids=["light", "fenced"])
def test_gbarrier_zz_benchmark(bfsmod, light):
    info = bfsmod.device_info()
    for grid in (info["num_cus"] // 4, info["num_cus"], 2 * info["num_cus"], 6 * info["num_cus"]):
        ms = bfsmod.gbarrier_benchmark(grid, 2000, light)
        print("gbarrier %.2f us/sync at grid %d light=%s" % (ms * 1e3, grid, light), flush=True)
        assert 0 < ms < 1.0


def test_from_host_problem_signature(bfsmod, orc):
    off, cols = orc.build_rmat_csr(10, 16, seed=4)
    prob = bfsmod.CsrProblem()
    assert prob.FromHostProblem(False, off.size - 1, cols.size, cols, off, 1) == 0
    en = bfsmod.EnactorHybrid()
    src = int(orc.pick_sources(off, 3, 1)[0])
    assert prob.Reset(en.GetFrontierType(), 1.3) == 0
    assert en.EnactSearch(prob, src) == 0
    labels = np.empty(off.size - 1, np.int32)
    assert prob.ExtractResults(labels) == 0
    np.testing.assert_array_equal(labels, orc.bfs(off, cols, src, with_pred=False)[0])
    assert prob.FromHostProblem(False, off.size - 1, cols.size, cols, off, 2) == b4._lib.ERR_UNSUPPORTED


@pytest.fixture(scope="module")
def rmat22(bfsmod, orc):
    """The headline graph at full size (RMAT scale 22, edge factor 16,
    a=0.57 b=0.19 c=0.19, seed 0) and all 64 of its seeded sources, with the
    oracle's depths for each."""
    g = bfsmod.Graph.rmat(22, 16, 0.57, 0.19, 0.19, 0)
    off, cols = g.to_host()
    srcs = [int(s) for s in orc.pick_sources(off, 1, 64)]
    refs = orc.bfs_many(off, cols, srcs, threads=16)
    return g, off, cols, srcs, refs


@pytest.mark.parametrize("strategy,mode,nsrc", [(b4.DIRECTION_OPTIMIZING, b4.MODE_DEFAULT, 64),
                                                (b4.EXPAND_CONTRACT, b4.MODE_ITERATIVE, 64),
                                                (b4.HYBRID, b4.MODE_DEFAULT, 64),
                                                (b4.EXPAND_CONTRACT, b4.MODE_FUSED, 8),
                                                (b4.CONTRACT_EXPAND, b4.MODE_FUSED, 8),
                                                (b4.TWO_PHASE, b4.MODE_FUSED, 8),
                                                (b4.TWO_PHASE, b4.MODE_ITERATIVE, 8)],
                         ids=["5-0", "0-2", "3-0", "0-1", "1-1", "2-1", "2-2"])
def test_rmat22_headline_sources(bfsmod, orc, rmat22, strategy, mode, nsrc):
    """Full-size parity on the headline workload, with the tuned knobs the
    bench uses: depth bit-exact, every predecessor valid, search depth.  The
    headline strategy and the reference's default hybrid run all 64 sources."""
    from back40computing_amd import autotune
    g, off, cols, srcs, refs = rmat22
    info = bfsmod.device_info()
    tuning = autotune.lookup(info["arch"], strategy, g.nodes, g.edges)
    preds = []
    for k, s in enumerate(srcs[:nsrc]):
        depth, pred, st = run(bfsmod, g, s, strategy, mode, tuning=tuning)
        ref = refs[k]
        sd = int(ref.max()) + 1
        np.testing.assert_array_equal(depth, ref)
        assert st["search_depth"] == sd
        preds.append(pred)
    # Predecessor checks on the host, several sources at a time.
    assert orc.check_pred_many(off, cols, refs[:nsrc], preds) == [-1] * nsrc


@pytest.fixture(scope="module")
def lattice200(bfsmod):
    """The high-diameter workload at full size: 200^3 6-neighbour lattice,
    ids permuted from seed 0 (V = 8,000,000, E = 47,760,000)."""
    g, perm = bfsmod.Graph.lattice3d(200, 0, want_perm=True)
    off, cols = g.to_host()
    return g, perm, off, cols


@pytest.mark.parametrize("strategy,mode,nsrc", [(b4.EXPAND_CONTRACT, b4.MODE_FUSED, 16),
                                                (b4.HYBRID, b4.MODE_DEFAULT, 16),
                                                (b4.DIRECTION_OPTIMIZING, b4.MODE_DEFAULT, 16),
                                                (b4.EXPAND_CONTRACT, b4.MODE_ITERATIVE, 4),
                                                (b4.CONTRACT_EXPAND, b4.MODE_FUSED, 4),
                                                (b4.TWO_PHASE, b4.MODE_FUSED, 4),
                                                (b4.TWO_PHASE, b4.MODE_ITERATIVE, 4)],
                         ids=["0-1", "3-0", "5-0", "0-2", "1-1", "2-1", "2-2"])
def test_lattice200_closed_form(bfsmod, orc, lattice200, strategy, mode, nsrc):
    """Full-size lattice: depth equals the Manhattan distance on the
    unpermuted coordinates (up to 598 levels from a corner).  Sources: two
    corners and a middle vertex, then the lattice workload's seeded sources
    (16 in all)."""
    g, perm, off, cols = lattice200
    assert g.nodes == 8000000 and g.edges == 47760000
    w = 200
    fixed = [int(perm[0]), int(perm[w ** 3 - 1]), int(perm[w ** 3 // 2 + 100])]
    seeded = [int(x) for x in orc.pick_sources(off, 3, 13)]
    refs, preds = [], []
    for s in (fixed + seeded)[:nsrc]:
        depth, pred, st = run(bfsmod, g, s, strategy, mode)
        ref = orc.lattice_depth(200, perm, s)
        np.testing.assert_array_equal(depth, ref)
        assert st["search_depth"] == int(ref.max()) + 1
        refs.append(ref)
        preds.append(pred)
    assert orc.check_pred_many(off, cols, refs, preds) == [-1] * len(refs)
