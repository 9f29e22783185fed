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
