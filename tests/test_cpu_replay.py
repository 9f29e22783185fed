"""CPU replay of the BFS kernel bodies under AddressSanitizer (no GPU).

tests/cpu_replay builds the library's kernel bodies (csrc/bfs_kernels.hpp)
with g++ against a host-thread emulation of the HIP device API: one host
thread per lane, wave collectives through a per-wave barrier, every device
buffer an allocation of exactly the library's size, and every workgroup's
LDS block filled with a garbage byte before the kernel starts (LDS keeps what
an earlier kernel left).  Depth labels are compared bit-exactly with the
serial oracle (or the lattice closed form), predecessors with the reference
invariants (test/bfs/test_bfs.cu:387-438).

The persistent two-phase kernel is the regression case: when its vertex
queue's count in LDS was aliased by the pair queue's values, stale LDS words
entered the vertex frontier and were used as addresses (a sanitizer SEGV with
0x3C fill, a heap overflow at offsets[-1] with 0xFF fill, wrong labels with
zero fill).
"""
import fcntl
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import csr_from_edges, hand_graphs

HERE = os.path.dirname(os.path.abspath(__file__))
REPLAY_DIR = os.path.join(HERE, "cpu_replay")
EXE = os.path.join(REPLAY_DIR, "replay_bfs")

EC, CE, TP, HYBRID, PERSISTENT = 0, 1, 2, 3, 4
DEFAULT, FUSED, ITERATIVE = 0, 1, 2


@pytest.fixture(scope="module")
def replay_exe():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    # One build at a time (parallel test workers share the output file).
    with open(os.path.join(REPLAY_DIR, ".build.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", REPLAY_DIR], check=True)
    return EXE


@pytest.fixture(scope="module")
def workdir():
    d = tempfile.mkdtemp(prefix="b40c_replay_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def replay(exe, workdir, off, cols, src, strategy, mode, grid=2, cta=256, wave=64, grid_th=2048, sat=512,
           queue_sizing=1.3, lds_fill=0x3C):
    g = os.path.join(workdir, "graph.bin")
    o = os.path.join(workdir, "out.bin")
    with open(g, "wb") as f:
        f.write(np.int32(off.size - 1).tobytes())
        f.write(np.int64(cols.size).tobytes())
        f.write(np.ascontiguousarray(off, np.int32).tobytes())
        f.write(np.ascontiguousarray(cols, np.int32).tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    cmd = [exe, g, str(src), str(strategy), str(mode), str(grid), str(cta), str(wave), str(grid_th), str(sat),
           str(queue_sizing), str(lds_fill), o]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, "replay failed (%s):\n%s" % (" ".join(cmd[2:-1]), r.stderr[-4000:])
    b = np.fromfile(o, np.int32)
    n = off.size - 1
    return int(b[0]), int(b[1]), b[2:2 + n], b[2 + n:2 + 2 * n]


def check(orc, off, cols, src, result, ref=None):
    status, sd, depth, pred = result
    assert status == 0
    if ref is None:
        ref = orc.bfs(off, cols, int(src), with_pred=False)[0]
    np.testing.assert_array_equal(depth, ref)
    assert orc.check_pred(off, cols, ref, pred) == -1
    assert sd == int(ref.max()) + 1


ALL = [(EC, ITERATIVE), (EC, FUSED), (CE, ITERATIVE), (CE, FUSED), (TP, ITERATIVE), (TP, FUSED), (HYBRID, DEFAULT),
       (PERSISTENT, DEFAULT)]


@pytest.fixture(scope="module")
def lattice8(orc):
    off, cols, perm = orc.build_lattice_csr(8, 5)
    return off, cols, perm


@pytest.mark.parametrize("sm", ALL, ids=["%d-%d" % sm for sm in ALL])
def test_replay_lattice_closed_form(replay_exe, workdir, orc, lattice8, sm):
    """Permuted 8^3 lattice from a corner: depth is the Manhattan distance.
    Hybrid runs with saturation_quit 2, so it hands off from the persistent
    contract-expand kernel to two-phase levels and back."""
    off, cols, perm = lattice8
    src = int(perm[0])
    res = replay(replay_exe, workdir, off, cols, src, *sm, grid=3, sat=2)
    check(orc, off, cols, src, res, ref=orc.lattice_depth(8, perm, src))


@pytest.mark.parametrize("fill", [0x00, 0x3C, 0xFF])
def test_replay_persistent_two_phase_any_lds(replay_exe, workdir, orc, lattice8, fill):
    """The persistent two-phase kernel is correct whatever LDS held before it
    started (the aliased queue count made it read stale LDS words)."""
    off, cols, perm = lattice8
    src = int(perm[8 ** 3 // 2 + 7])
    res = replay(replay_exe, workdir, off, cols, src, TP, FUSED, grid=3, lds_fill=fill)
    check(orc, off, cols, src, res, ref=orc.lattice_depth(8, perm, src))


@pytest.fixture(scope="module")
def rmat10(orc):
    off, cols = orc.build_rmat_csr(10, 16, seed=3)
    src = int(orc.pick_sources(off, 1, 1)[0])
    return off, cols, src


@pytest.mark.parametrize("sm", ALL, ids=["%d-%d" % sm for sm in ALL])
def test_replay_rmat_gather_tiers(replay_exe, workdir, orc, rmat10, sm):
    """RMAT-10 with low gather thresholds, so the workgroup, wave and fine
    tiers (and, for the level kernels, the grid tier) all run."""
    off, cols, src = rmat10
    res = replay(replay_exe, workdir, off, cols, src, *sm, grid=2, cta=96, wave=8, grid_th=200, sat=2)
    check(orc, off, cols, src, res)


def test_replay_hand_graphs_persistent(replay_exe, workdir, orc):
    """Hand-worked graphs (paths, stars, cliques, pieces, isolated sources)
    through the three persistent kernels."""
    for g in hand_graphs():
        off, cols = csr_from_edges(g["n"], g["edges"], g["undirected"])
        for case in g["cases"]:
            for sm in [(EC, FUSED), (CE, FUSED), (TP, FUSED)]:
                status, sd, depth, pred = replay(replay_exe, workdir, off, cols, case["src"], *sm, grid=2)
                assert status == 0
                assert depth.tolist() == case["depth"], (g["name"], sm)
                assert sd == case["search_depth"]
                assert orc.check_pred(off, cols, depth, pred) == -1


@pytest.mark.parametrize("sm", [(CE, FUSED), (TP, FUSED), (EC, FUSED), (TP, ITERATIVE)],
                         ids=["1-1", "2-1", "0-1", "2-2"])
def test_replay_forced_overflow(replay_exe, workdir, rmat10, sm):
    """A tiny queue-sizing factor ends in the overflow status, with every
    access still inside its buffer."""
    off, cols, src = rmat10
    status, _, _, _ = replay(replay_exe, workdir, off, cols, src, *sm, grid=2, queue_sizing=0.01)
    assert status == 1


@pytest.mark.parametrize("grid", [2, 5])
@pytest.mark.parametrize("light", [1, 0])
def test_replay_grid_barrier(replay_exe, grid, light):
    """The grid barrier's visibility check (words written before a barrier
    are seen after it, and the payload workgroup 0 publishes is the same in
    every workgroup), over two launches that continue one epoch sequence,
    in both barrier modes."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([replay_exe, "--gbarrier", str(grid), "30", str(light)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.split() == ["mismatches", "0", "aborted", "0", "epoch", str(2 * 30 * 2)]
