"""The autotuner on the GPU: a short coordinate descent on RMAT-12 for two
strategies, recorded into a scratch tuning table; the recorded entry is
looked up again and its labels are compared with the serial oracle on the
tuning sources and on sources it never saw."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("strategy,knobs", [(5, ["do_alpha", "grid_gather_threshold"]),
                                            (3, ["saturation_quit", "cta_gather_threshold"])])
def test_tuned_entry_reproduces_oracle_labels(orc, tmp_path, strategy, knobs):
    from back40computing_amd import autotune, bfs
    g = bfs.Graph.rmat(12, 16, 0.57, 0.19, 0.19, 2)
    off, cols = g.to_host()
    srcs = [int(s) for s in orc.pick_sources(off, 7, 4)]
    path = str(tmp_path / "table.json")
    print("autotune strategy=%d knobs=%s" % (strategy, knobs), flush=True)
    best, ms, hist = autotune.tune(g, strategy, srcs, rounds=1, repeats=1, knobs=knobs)
    assert ms > 0 and len(hist) > 1
    assert all(h[1] is not None for h in hist), "a candidate gave labels that differ from the verified ones"
    table = autotune.load_table(path)
    assert table == {}
    best2, _ = autotune.tune_and_record(g, strategy, srcs, path=path, rounds=1, repeats=1, knobs=knobs)
    info = bfs.device_info()
    entry = autotune.lookup(info["arch"], strategy, g.nodes, g.edges, path=path)
    assert entry == best2
    unseen = [int(s) for s in orc.pick_sources(off, 99, 4)]
    for s in srcs + unseen:
        depth, pred, st = bfs.bfs(g, s, strategy=strategy, tuning=entry)
        ref, _, sd = orc.bfs(off, cols, s, with_pred=False)
        np.testing.assert_array_equal(depth, ref)
        assert orc.check_pred(off, cols, ref, pred) == -1
        assert st["search_depth"] == sd
    g.close()
