"""Robustness of the graph-file parsers (host code reading untrusted text):
mutated and random files must give either an error status or a well-formed
edge list (ids in range, offsets monotone) — never a crash."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "graph_files")


@pytest.fixture(scope="module")
def gio():
    from back40computing_amd import _lib, graph_io
    if not os.path.exists(_lib.LIB_PATH):
        from back40computing_amd import build
        build.build()
    return graph_io


ALPHABET = np.frombuffer(b"0123456789 \t\n\r%-+apce.x", np.uint8)


def mutate(rng, data):
    b = bytearray(data)
    for _ in range(int(rng.integers(1, 6))):
        op = rng.integers(0, 4)
        if op == 0 and b:  # flip a byte to a format-relevant character
            b[int(rng.integers(0, len(b)))] = int(rng.choice(ALPHABET))
        elif op == 1:  # insert
            pos = int(rng.integers(0, len(b) + 1))
            b[pos:pos] = bytes(rng.choice(ALPHABET, int(rng.integers(1, 8))).tolist())
        elif op == 2 and b:  # delete a run
            pos = int(rng.integers(0, len(b)))
            del b[pos:pos + int(rng.integers(1, 8))]
        else:  # truncate
            b = b[:int(rng.integers(0, len(b) + 1))]
    return bytes(b)


def check_well_formed(el):
    a, b = el.arrays()
    n = el.nodes
    assert n >= 1
    assert np.all((b >= 0) & (b < n))
    if el.is_csr:
        assert a.size == n + 1 and a[0] == 0 and a[-1] == b.size
        assert np.all(np.diff(a) >= 0)
    else:
        assert a.size == b.size and np.all((a >= 0) & (a < n))


@pytest.mark.parametrize("fmt,name", [("market", "tiny.mtx"), ("metis", "tiny.graph"), ("metis", "comment_joins.graph"),
                                      ("dimacs", "tiny.gr")])
def test_mutated_fixtures(gio, tmp_path, fmt, name):
    from back40computing_amd import B40CError
    base = open(os.path.join(FIX, name), "rb").read()
    rng = np.random.default_rng(sum(name.encode()))
    path = str(tmp_path / name)
    ok = err = 0
    for i in range(400):
        data = mutate(rng, base)
        if re.search(rb"\d{8}", data):
            continue  # a header naming 10^7+ vertices is legal but allocates accordingly
        open(path, "wb").write(data)
        for undirected in (0, 1):
            try:
                el = gio.EdgeList(path, fmt, undirected)
            except B40CError:
                err += 1
                continue
            check_well_formed(el)
            el.close()
            ok += 1
    assert ok > 0 and err > 0  # both outcomes are exercised


@pytest.mark.parametrize("fmt", ["market", "metis", "dimacs"])
def test_random_bytes(gio, tmp_path, fmt):
    from back40computing_amd import B40CError
    rng = np.random.default_rng(len(fmt))
    path = str(tmp_path / "noise")
    for i in range(300):
        data = bytes(rng.choice(ALPHABET, int(rng.integers(0, 400))).tolist())
        if i % 3 == 0:  # a plausible header in front of the noise
            data = {"market": b"5 5 3\n", "metis": b"5 3\n", "dimacs": b"p sp 5 3\n"}[fmt] + data
        open(path, "wb").write(data)
        try:
            el = gio.EdgeList(path, fmt, i % 2)
        except B40CError:
            continue
        check_well_formed(el)
        el.close()
