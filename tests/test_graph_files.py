"""Graph-file readers (MatrixMarket, METIS, DIMACS) on the host: the
library's parse (b40c_edge_list_*, no GPU) against the hand-worked fixtures,
the oracle's restatement of the reference readers, and files written from a
known edge list (whose reference edge order is fixed by construction)."""
import json
import os

import numpy as np
import pytest

from oracle import graph_files as orc_files

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "graph_files")
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "graph_files.json")))


@pytest.fixture(scope="module")
def gio():
    from back40computing_amd import _lib, graph_io
    if not os.path.exists(_lib.LIB_PATH):
        from back40computing_amd import build
        build.build()
    return graph_io


def lib_csr(el):
    a, b = el.arrays()
    if el.is_csr:
        return el.nodes, a, b
    return orc_files.from_coo(el.nodes, a, b)


def case_path(c):
    return os.path.join(FIX, c["file"])


@pytest.mark.parametrize("c", GOLDEN["cases"], ids=lambda c: "%s-u%d" % (c["file"], c["undirected"]))
def test_fixtures(gio, c):
    el = gio.EdgeList(case_path(c), c["format"], c["undirected"], c["splice"])
    n, off, cols = lib_csr(el)
    assert n == c["nodes"]
    assert off.tolist() == c["offsets"]
    assert cols.tolist() == c["cols"]


@pytest.mark.parametrize("c", GOLDEN["errors"], ids=lambda c: c["file"])
def test_fixture_errors(gio, c):
    from back40computing_amd import B40CError
    with pytest.raises(B40CError):
        gio.EdgeList(case_path(c), c["format"], c["undirected"])


def test_missing_file_and_bad_arguments(gio, tmp_path):
    from back40computing_amd import B40CError, _lib
    with pytest.raises(B40CError) as ei:
        gio.EdgeList(str(tmp_path / "none.mtx"), "market")
    assert ei.value.status == _lib.ERR_INVALID_ARG
    with pytest.raises(B40CError):
        gio.EdgeList(os.path.join(FIX, "tiny.mtx"), 7)
    with pytest.raises(B40CError):  # splice is a DIMACS option
        gio.EdgeList(os.path.join(FIX, "tiny.mtx"), "market", splice=2)
    with pytest.raises(B40CError) as ei:
        gio.EdgeList(os.path.join(FIX, "weighted.graph"), "metis")
    assert ei.value.status == _lib.ERR_UNSUPPORTED


# ---- files written from a known edge list --------------------------------

def random_entries(rng, n, m):
    """1-based (i, j) pairs with self-loops and duplicates."""
    i = rng.integers(1, n + 1, m)
    j = rng.integers(1, n + 1, m)
    dup = rng.integers(0, m, m // 10)
    i[dup[: len(dup) // 2]] = i[dup[len(dup) // 2:][: len(dup) // 2]]
    j[:5] = i[:5]  # self-loops
    return i, j


def write_market(path, n, i, j, rng):
    lines = ["%%MatrixMarket matrix coordinate pattern general", "% written by the test"]
    lines.append("%d %d %d" % (n, n, len(i)))
    for k in range(len(i)):
        r = rng.random()
        if r < 0.02:
            lines.append("% comment")
        elif r < 0.04:
            lines.append("")
        pre = "  " if rng.random() < 0.05 else ""
        tail = " %.3f" % rng.random() if rng.random() < 0.3 else ""
        sep = "\t" if rng.random() < 0.1 else " "
        lines.append("%s%d%s%d%s" % (pre, i[k], sep, j[k], tail))
    open(path, "w").write("\n".join(lines) + "\n")


def market_coo(i, j, undirected):
    """Reference order: entry (i, j) is edge j-1 -> i-1, reverse right after."""
    if not undirected:
        return j - 1, i - 1
    rows = np.empty(2 * len(i), np.int64)
    cols = np.empty(2 * len(i), np.int64)
    rows[0::2], cols[0::2] = j - 1, i - 1
    rows[1::2], cols[1::2] = i - 1, j - 1
    return rows, cols


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("undirected", [0, 1])
def test_market_random(gio, tmp_path, seed, undirected):
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(2, 3000)), int(rng.integers(1, 20000))
    i, j = random_entries(rng, n, m)
    path = str(tmp_path / "g.mtx")
    write_market(path, n, i, j, rng)
    el = gio.EdgeList(path, "market", undirected)
    rows, cols = el.arrays()
    er, ec = market_coo(i, j, undirected)
    np.testing.assert_array_equal(rows, er)
    np.testing.assert_array_equal(cols, ec)
    on, orow, ocol = orc_files.read_market_coo(open(path).read(), undirected)
    assert on == n
    np.testing.assert_array_equal(rows, orow)
    np.testing.assert_array_equal(cols, ocol)


def test_market_large_multithreaded(gio, tmp_path):
    """A file of several MiB is parsed in parallel chunks; the edge order
    must still be the file order."""
    rng = np.random.default_rng(9)
    n, m = 1 << 20, 700_000
    i = rng.integers(1, n + 1, m)
    j = rng.integers(1, n + 1, m)
    path = str(tmp_path / "big.mtx")
    body = "\n".join("%d %d" % (a, b) for a, b in zip(i.tolist(), j.tolist()))
    open(path, "w").write("%%%%MatrixMarket matrix coordinate pattern general\n%d %d %d\n%s\n" % (n, n, m, body))
    assert os.path.getsize(path) > 8 << 20
    el = gio.EdgeList(path, "market", 1)
    rows, cols = el.arrays()
    er, ec = market_coo(i, j, 1)
    np.testing.assert_array_equal(rows, er)
    np.testing.assert_array_equal(cols, ec)


def write_metis(path, n, adj, m_header, rng, drop_trailing):
    lines = ["% METIS file written by the test", "%d %d" % (n, m_header)]
    last = n
    if drop_trailing:
        while last > 0 and not adj[last - 1]:
            last -= 1
    for v in range(last):
        if rng.random() < 0.03:
            lines.append("% a comment line")
        sep = "\t" if rng.random() < 0.1 else " "
        lines.append(("  " if rng.random() < 0.05 else "") + sep.join(str(x + 1) for x in adj[v]))
    open(path, "w").write("\n".join(lines) + "\n")


@pytest.mark.parametrize("seed", [4, 5, 6])
def test_metis_random(gio, tmp_path, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 2000))
    adj = [sorted(rng.integers(0, n, int(rng.integers(0, 12))).tolist()) if rng.random() < 0.8 else []
           for _ in range(n)]
    for k in range(1, 6):
        adj[-k] = []  # trailing isolated vertices
    total = sum(len(a) for a in adj)
    m_header = total // 2 if total % 2 == 0 and rng.random() < 0.5 else total
    path = str(tmp_path / "g.graph")
    write_metis(path, n, adj, m_header, rng, drop_trailing=bool(seed % 2))
    el = gio.EdgeList(path, "metis")
    assert el.is_csr and el.nodes == n and el.edges == total
    off, cols = el.arrays()
    exp_off = np.concatenate([[0], np.cumsum([len(a) for a in adj])])
    np.testing.assert_array_equal(off, exp_off)
    np.testing.assert_array_equal(cols, [x for a in adj for x in a])
    on, ooff, ocol = orc_files.read_metis(open(path).read())
    np.testing.assert_array_equal(off, ooff)
    np.testing.assert_array_equal(cols, ocol)


def write_dimacs(path, n, u, v, rng, m_header=None):
    lines = ["c DIMACS file written by the test", "p sp %d %d" % (n, len(u) if m_header is None else m_header)]
    for k in range(len(u)):
        if rng.random() < 0.03:
            lines.append("c comment")
        lines.append("a %d %d %d" % (u[k], v[k], int(rng.integers(1, 100))))
    open(path, "w").write("\n".join(lines) + "\n")


@pytest.mark.parametrize("seed", [7, 8])
@pytest.mark.parametrize("undirected", [0, 1])
def test_dimacs_random(gio, tmp_path, seed, undirected):
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(2, 3000)), int(rng.integers(1, 15000))
    u, v = random_entries(rng, n, m)
    path = str(tmp_path / "g.gr")
    write_dimacs(path, n, u, v, rng)
    el = gio.EdgeList(path, "dimacs", undirected)
    rows, cols = el.arrays()
    if undirected:  # reverse edge first, then the edge itself
        er = np.empty(2 * m, np.int64)
        ec = np.empty(2 * m, np.int64)
        er[0::2], ec[0::2] = v - 1, u - 1
        er[1::2], ec[1::2] = u - 1, v - 1
    else:
        er, ec = u - 1, v - 1
    np.testing.assert_array_equal(rows, er)
    np.testing.assert_array_equal(cols, ec)
    on, orow, ocol = orc_files.read_dimacs_coo([open(path).read()], undirected)
    np.testing.assert_array_equal(rows, orow)
    np.testing.assert_array_equal(cols, ocol)


@pytest.mark.parametrize("undirected", [0, 1])
def test_dimacs_splice(gio, tmp_path, undirected):
    """Splice files merged by row.  With each file's rows non-decreasing, the
    directed merge is the sort by (row, file, position)."""
    rng = np.random.default_rng(11)
    n, m, files = 500, 6000, 3
    u, v = random_entries(rng, n, m)
    order = np.argsort(u, kind="stable")
    u, v = u[order], v[order]
    which = rng.integers(0, files, m)
    base = str(tmp_path / "s.gr")
    for f in range(files):
        sel = which == f
        write_dimacs(base + ".%d" % f, n, u[sel], v[sel], rng)
    el = gio.EdgeList(base, "dimacs", undirected, files)
    rows, cols = el.arrays()
    texts = [open(base + ".%d" % f).read() for f in range(files)]
    on, orow, ocol = orc_files.read_dimacs_coo(texts, undirected)
    np.testing.assert_array_equal(rows, orow)
    np.testing.assert_array_equal(cols, ocol)
    if not undirected:
        pos = np.concatenate([np.arange(np.sum(which == f)) for f in range(files)])
        fidx = np.concatenate([np.full(np.sum(which == f), f) for f in range(files)])
        uu = np.concatenate([u[which == f] for f in range(files)])
        vv = np.concatenate([v[which == f] for f in range(files)])
        k = np.lexsort((pos, fidx, uu))
        np.testing.assert_array_equal(rows, uu[k] - 1)
        np.testing.assert_array_equal(cols, vv[k] - 1)
