"""GPU parity of graph files: the device CSR built from MatrixMarket, METIS
and DIMACS files (b40c_graph_read / b40c_graph_from_edge_list) against the
hand-worked fixtures and the oracle, METIS write -> read round trips, and
BFS on file-loaded graphs against the serial oracle."""
import json
import os

import numpy as np
import pytest

from oracle import graph_files as orc_files

pytestmark = pytest.mark.gpu

import back40computing_amd as b4  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "graph_files")
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "graph_files.json")))


@pytest.fixture(scope="module")
def mods():
    from back40computing_amd import bfs, graph_io
    return bfs, graph_io


def check_bfs(bfs, orc, g, off, cols, sources, strategies):
    for s in sources:
        ref, _, rsd = orc.bfs(off, cols, int(s), with_pred=False)
        for strat in strategies:
            print("file graph bfs strategy=%d src=%d n=%d" % (strat, s, g.nodes), flush=True)
            depth, pred, st = bfs.bfs(g, int(s), strategy=strat)
            np.testing.assert_array_equal(depth, ref)
            assert orc.check_pred(off, cols, ref, pred) == -1
            assert st["search_depth"] == rsd


@pytest.mark.parametrize("c", GOLDEN["cases"], ids=lambda c: "%s-u%d" % (c["file"], c["undirected"]))
def test_device_csr_of_fixtures(mods, c):
    bfs, gio = mods
    g = gio.read_graph(os.path.join(FIX, c["file"]), c["format"], c["undirected"], c["splice"])
    off, cols = g.to_host()
    assert g.nodes == c["nodes"]
    assert off.tolist() == c["offsets"]
    assert cols.tolist() == c["cols"]


def test_metis_round_trip_and_bfs(mods, orc, tmp_path):
    bfs, gio = mods
    g = bfs.Graph.rmat(12, 16, 0.57, 0.19, 0.19, 3)
    off, cols = g.to_host()
    path = str(tmp_path / "rmat12.graph")
    gio.write_metis(g, path)
    with open(path) as f:
        assert f.readline().split() == [str(g.nodes), str(g.edges)]
    g2 = gio.read_graph(path, "metis")
    off2, cols2 = g2.to_host()
    np.testing.assert_array_equal(off2, off)
    np.testing.assert_array_equal(cols2, cols)
    srcs = orc.pick_sources(off, 17, 3)
    check_bfs(bfs, orc, g2, off, cols, srcs, [b4.DIRECTION_OPTIMIZING, b4.EXPAND_CONTRACT])


def test_market_file_is_the_graph(mods, orc, tmp_path):
    """A MatrixMarket file listing a CSR's edges (entry "c+1 r+1" for edge
    r -> c, rows in order) reads back to the same CSR: the stable sort by row
    keeps each row's columns in file order."""
    bfs, gio = mods
    g = bfs.Graph.rmat(13, 16, 0.57, 0.19, 0.19, 5)
    off, cols = g.to_host()
    rows = np.repeat(np.arange(g.nodes), np.diff(off))
    path = str(tmp_path / "rmat13.mtx")
    with open(path, "w") as f:
        f.write("%%%%MatrixMarket matrix coordinate pattern general\n%d %d %d\n" % (g.nodes, g.nodes, g.edges))
        f.write("\n".join("%d %d" % (c + 1, r + 1) for r, c in zip(rows.tolist(), cols.tolist())))
        f.write("\n")
    g2 = gio.read_graph(path, "market", undirected=False)
    off2, cols2 = g2.to_host()
    np.testing.assert_array_equal(off2, off)
    np.testing.assert_array_equal(cols2, cols)
    check_bfs(bfs, orc, g2, off, cols, orc.pick_sources(off, 23, 2), [b4.HYBRID, b4.DIRECTION_OPTIMIZING])


@pytest.mark.parametrize("undirected", [0, 1])
def test_dimacs_file_bfs(mods, orc, tmp_path, undirected):
    """Random arcs with duplicates and self-loops: the device CSR equals the
    oracle's reading of the file, and BFS on it equals the serial BFS."""
    bfs, gio = mods
    rng = np.random.default_rng(40 + undirected)
    n, m = 5000, 40000
    u = rng.integers(1, n + 1, m)
    v = rng.integers(1, n + 1, m)
    v[:20] = u[:20]
    path = str(tmp_path / "arcs.gr")
    with open(path, "w") as f:
        f.write("c random arcs\np sp %d %d\n" % (n, m))
        f.write("\n".join("a %d %d %d" % (a, b, (a * 7 + b) % 100) for a, b in zip(u.tolist(), v.tolist())))
        f.write("\n")
    g = gio.read_graph(path, "dimacs", undirected=undirected)
    off, cols = g.to_host()
    _, eoff, ecols = orc_files.read_dimacs([open(path).read()], undirected)
    np.testing.assert_array_equal(off, eoff)
    np.testing.assert_array_equal(cols, ecols)
    strategies = [b4.EXPAND_CONTRACT, b4.DIRECTION_OPTIMIZING] if undirected else [b4.EXPAND_CONTRACT, b4.HYBRID]
    check_bfs(bfs, orc, g, off, cols, orc.pick_sources(off, 29, 2), strategies)
