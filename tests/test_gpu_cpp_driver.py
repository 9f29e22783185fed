"""The C++ parity driver (tests/cpp/test_bfs.cpp) on the GPU: it reaches the
library only through the C++ facade (include/b40c/bfs.hpp), the way a
reference-side caller would, and validates every source against the CPU
oracle and the predecessor invariants (test/bfs/test_bfs.cu:387-438)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_bfs")

CASES = [
    ("rmat", "12", "--strategy=0"),
    ("rmat", "12", "--strategy=1"),
    ("rmat", "12", "--strategy=2"),
    ("rmat", "12", "--strategy=3"),
    ("rmat", "12", "--strategy=5"),
    ("grid3d", "20", "--strategy=0"),
    ("grid3d", "20", "--strategy=3"),
]


@pytest.mark.parametrize("case", CASES, ids=["-".join(c).replace("--strategy=", "s") for c in CASES])
def test_cpp_driver(case):
    if not os.path.exists(DRIVER):
        pytest.fail("C++ driver not built (run __graft_entry__.build())")
    print("driver case:", case, flush=True)
    r = subprocess.run([DRIVER, *case, "--i=3", "--seed=5"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures=0" in r.stdout
    assert r.stdout.count("CORRECT") >= 3


def test_cpp_sort_driver():
    """The radix sort through the C++ facade (include/b40c/radix_sort.hpp):
    64-bit keys (all bits and a bit range) and stable 32-bit pairs against
    std::sort / std::stable_sort (test/radix_sort/test_sort.cu:185-261)."""
    drv = os.path.join(os.path.dirname(DRIVER), "test_sort")
    if not os.path.exists(drv):
        pytest.fail("C++ sort driver not built (run __graft_entry__.build())")
    r = subprocess.run([drv, "3000000", "11"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all CORRECT" in r.stdout


def _grid_edges(w):
    """Undirected edges (1-based) of a w x w grid, each listed once."""
    out = []
    for r in range(w):
        for c in range(w):
            v = r * w + c + 1
            if c + 1 < w:
                out.append((v, v + 1))
            if r + 1 < w:
                out.append((v, v + w))
    return out


@pytest.mark.parametrize("kind", ["market", "metis", "dimacs"])
def test_cpp_driver_file_graphs(tmp_path, kind):
    """The driver's file graph types (test_bfs.cu:976-1037) through the
    facade's builders (include/b40c/graph_builder.hpp)."""
    if not os.path.exists(DRIVER):
        pytest.fail("C++ driver not built (run __graft_entry__.build())")
    w = 60
    n, edges = w * w, _grid_edges(w)
    path = str(tmp_path / ("grid." + kind))
    with open(path, "w") as f:
        if kind == "market":
            f.write("%%%%MatrixMarket matrix coordinate pattern symmetric\n%d %d %d\n" % (n, n, len(edges)))
            f.write("".join("%d %d\n" % e for e in edges))
        elif kind == "dimacs":
            f.write("p sp %d %d\n" % (n, len(edges)))
            f.write("".join("a %d %d 1\n" % e for e in edges))
        else:
            adj = [[] for _ in range(n)]
            for a, b in edges:
                adj[a - 1].append(b)
                adj[b - 1].append(a)
            f.write("%d %d\n" % (n, len(edges)))
            f.write("".join(" ".join(map(str, sorted(a))) + "\n" for a in adj))
    extra = [] if kind == "metis" else ["--undirected"]
    for strategy in ("--strategy=0", "--strategy=5"):
        print("driver file case:", kind, strategy, flush=True)
        r = subprocess.run([DRIVER, kind, path, *extra, strategy, "--i=3", "--seed=2"], capture_output=True,
                           text=True, timeout=120)
        print(r.stdout[-1500:])
        assert r.returncode == 0, r.stdout + r.stderr
        assert "graph %s: nodes=%d edges=%d" % (kind, n, 2 * len(edges)) in r.stdout
        assert "failures=0" in r.stdout


@pytest.mark.parametrize("strategy", ["--strategy=0", "--strategy=2", "--strategy=3"])
def test_cpp_driver_instrumented_histogram(strategy):
    """--instrumented runs the instrumented enactors (average workgroup duty
    reported), every source line carries total queued and redundant work, the
    summary the bias-corrected statistics, and --v2 the work histogram by
    depth (test_bfs.cu:246-337, 343-490)."""
    if not os.path.exists(DRIVER):
        pytest.fail("C++ driver not built (run __graft_entry__.build())")
    print("driver instrumented case:", strategy, flush=True)
    r = subprocess.run([DRIVER, "rmat", "12", strategy, "--i=3", "--seed=4", "--instrumented", "--v2"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2500:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures=0" in r.stdout and "instrumented=1" in r.stdout
    assert r.stdout.count("Work Histogram:") == 3
    assert "avg_duty=" in r.stdout and "redundant_work=" in r.stdout
    assert "Summary after 3 test iterations" in r.stdout and "[Duty %]" in r.stdout
    # The histogram's discovered column sums to the nodes visited of its source.
    lines = r.stdout.splitlines()
    i = lines.index("Depth, Expanded, Unique-Expanded, Discovered")
    src_line = lines[i - 2]
    nv = int(src_line.split("nodes_visited=")[1].split()[0])
    disc = 0
    for ln in lines[i + 1:]:
        parts = ln.split(", ")
        if len(parts) != 4:
            break
        disc += int(parts[3])
    assert disc == nv


MB_CASES = [("rmat", "12", "--microbench=gather_lookup"), ("rmat", "12", "--microbench=cull"),
            ("grid3d", "16", "--microbench=cull")]


@pytest.mark.parametrize("case", MB_CASES, ids=["-".join(c).replace("--microbench=", "") for c in MB_CASES])
def test_cpp_driver_microbench(case):
    """--microbench: the reference's microbench_bfs / microbench_cull drivers
    (the test_bfs loop over the per-stage microbenchmark search): per-iteration
    queue table, labels CORRECT against the CPU BFS, culling efficiency in
    [0, 100] % (microbench_cull.cu:313-331)."""
    if not os.path.exists(DRIVER):
        pytest.fail("C++ driver not built (run __graft_entry__.build())")
    print("driver microbench case:", case, flush=True)
    r = subprocess.run([DRIVER, *case, "--i=3", "--seed=6"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2500:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures=0" in r.stdout and r.stdout.count(": CORRECT") == 3
    assert r.stdout.count("Compaction queue, Expansion queue") == 3
    effs = [float(ln.split("culling_efficiency=")[1].split("%")[0]) for ln in r.stdout.splitlines()
            if ln.startswith("src ")]
    assert len(effs) == 3 and all(0.0 <= e <= 100.0 for e in effs)
    if "cull" in case[-1] and "gather" not in case[-1]:
        assert min(effs) > 0.0
