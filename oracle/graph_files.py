"""CPU oracle for the graph-file readers — test infrastructure only.

Pure-Python restatement of the reference's text readers, stream semantics
included, for small files:

* ``read_market``  — ReadMarketStream, b40c/graph/builder/market.cuh:50-158
* ``read_metis``   — ReadMetisStream,  b40c/graph/builder/metis.cuh:52-158
* ``read_dimacs``  — ReadDimacsStream, b40c/graph/builder/dimacs.cuh:54-220
* ``from_coo``     — CsrGraph::FromCoo, b40c/graph/csr_graph.cuh:108-152

Each returns ``(nodes, row_offsets, cols)`` as numpy int32 arrays or raises
``ValueError`` where the reference reports an error.  Where the reference
would write out of bounds or build a corrupt CSR (ids outside [1, n], more
METIS vertex lines than vertices) this restatement raises instead.  Only
``tests/`` may import it.
"""
import numpy as np

_WS = " \t\n\r\v\f"


def _fscanf_lines(text):
    """Lines as the reference's ``fscanf(f, "%[^\\n]\\n", line)`` loop
    yields them (market.cuh:71, dimacs.cuh:74,125): the conversion needs at
    least one non-newline character (so the loop ends if the stream starts
    with a newline) and the trailing "\\n" directive skips all whitespace,
    i.e. blank lines and the next line's leading blanks."""
    out, i, n = [], 0, len(text)
    while i < n:
        if text[i] == "\n":
            break
        j = text.find("\n", i)
        j = n if j < 0 else j
        out.append(text[i:j])
        i = j
        while i < n and text[i] in _WS:
            i += 1
    return out


def _scan_ints(s, count):
    """``sscanf(s, "%lld %lld ...")``: up to `count` integers, stopping at the
    first that does not convert; returns the list converted."""
    vals, i, n = [], 0, len(s)
    while len(vals) < count:
        while i < n and s[i] in _WS:
            i += 1
        j = i
        if j < n and s[j] in "+-":
            j += 1
        k = j
        while k < n and s[k].isdigit():
            k += 1
        if k == j:
            break
        vals.append(int(s[i:k]))
        i = k
    return vals


def from_coo(nodes, rows, cols):
    """Stable sort by row only, then offsets (csr_graph.cuh:108-152)."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    order = np.argsort(rows, kind="stable")
    off = np.zeros(nodes + 1, np.int64)
    np.add.at(off, rows + 1, 1)
    return nodes, np.cumsum(off).astype(np.int32), cols[order].astype(np.int32)


def _check_id(x, n):
    if x < 1 or x > n:
        raise ValueError("vertex id %d outside [1, %d]" % (x, n))


def read_market(text, undirected):
    """CSR of a MatrixMarket text (read_market_coo, then from_coo)."""
    return from_coo(*read_market_coo(text, undirected))


def read_market_coo(text, undirected):
    """(nodes, rows, cols) in file order; market.cuh:50-158.  Entry "i j" is the edge (j-1) -> (i-1)
    (``sscanf(line, "%lld %lld", &ll_col, &ll_row)``, :108); undirected
    appends the reverse edge right after it (:117-124)."""
    nodes = edges = None
    rows, cols = [], []
    for line in _fscanf_lines(text):
        if line[0] == "%":
            continue
        if nodes is None:
            v = _scan_ints(line, 3)
            if len(v) != 3:
                raise ValueError("invalid problem description")
            if v[0] != v[1]:
                raise ValueError("not square")
            nodes, edges = v[0], (2 * v[2] if undirected else v[2])
            continue
        if len(rows) >= edges:
            raise ValueError("more edges than declared")
        v = _scan_ints(line, 2)
        if len(v) != 2:
            raise ValueError("badly formed edge")
        c, r = v
        _check_id(c, nodes)
        _check_id(r, nodes)
        rows.append(r - 1)
        cols.append(c - 1)
        if undirected:
            rows.append(c - 1)
            cols.append(r - 1)
    if nodes is None:
        raise ValueError("no graph found")
    if len(rows) != edges:
        raise ValueError("only %d/%d edges read" % (len(rows), edges))
    return nodes, rows, cols


def read_metis(text):
    """metis.cuh:52-158, character by character: '%' skips through its
    newline (so a comment line is not a vertex line), ' ' and '\\t' are
    skipped, '\\n' starts the next vertex, anything else is read with
    ``fscanf("%lld")`` — the header "n m ..." first (the rest of its line
    ignored), then 1-based neighbour ids."""
    i, n_chars = 0, len(text)
    nodes = m2 = None
    current = -1
    offsets = None
    cols = []
    while i < n_chars:
        c = text[i]
        i += 1
        if c == "%":
            j = text.find("\n", i)
            i = n_chars if j < 0 else j + 1
        elif c in " \t":
            pass
        elif c == "\n":
            if offsets is None:
                raise ValueError("newline before the header")  # the reference writes a NULL array here
            current += 1
            if current <= nodes:  # later (blank) lines: ids there raise below
                offsets[current] = len(cols)
        else:
            i -= 1
            if current == -1:
                j = text.find("\n", i)
                j = n_chars if j < 0 else j
                v = _scan_ints(text[i:j], 2)
                if len(v) < 2:
                    raise ValueError("invalid format")
                nodes, m2 = v[0], 2 * v[1]
                offsets = np.zeros(nodes + 1, np.int64)
                i = j  # "%[^\n]" stops before the newline
                continue
            # fscanf("%lld") skips leading whitespace, newlines included
            j = i
            while j < n_chars and text[j] in _WS:
                j += 1
            k = j + 1 if j < n_chars and text[j] in "+-" else j
            e = k
            while e < n_chars and text[e].isdigit():
                e += 1
            if e == k:
                raise ValueError("invalid format")
            x = int(text[j:e])
            i = e
            if len(cols) >= m2:
                raise ValueError("more edges than declared")
            _check_id(x, nodes)
            if current >= nodes:
                raise ValueError("more vertex lines than vertices")
            cols.append(x - 1)
    if offsets is None:
        raise ValueError("no graph found")
    while current < nodes:
        current += 1
        offsets[current] = len(cols)
    if m2 == 2 * len(cols):
        pass  # the header gave the directed count (:141-145)
    elif len(cols) != m2:
        raise ValueError("only %d/%d edges read" % (len(cols), m2))
    return nodes, offsets.astype(np.int32), np.asarray(cols, np.int32)


def read_dimacs(texts, undirected):
    """CSR of one DIMACS text or a splice (read_dimacs_coo, then from_coo)."""
    return from_coo(*read_dimacs_coo(texts, undirected))


def read_dimacs_coo(texts, undirected):
    """(nodes, rows, cols) in the reference's order; dimacs.cuh:54-220 over one text or a splice of several: each file's
    lines up to its "p <type> n m" line are skipped; afterwards "a u v w"
    lines are edges.  Edges are merged by row across files (strict <, so
    ties go to the lowest file), and in undirected mode an edge's reverse is
    emitted as soon as its line is read (:127-147)."""
    lines = [_fscanf_lines(t) for t in texts]
    pos = [0] * len(texts)
    nodes = edges = 0
    for f, ls in enumerate(lines):
        while pos[f] < len(ls):
            line = ls[pos[f]]
            pos[f] += 1
            if line[0] == "p":
                parts = line.split()
                if len(parts) < 4:
                    raise ValueError("invalid problem line")
                n, m = int(parts[2]), int(parts[3])
                if nodes and nodes != n:
                    raise ValueError("splice files do not name the same number of vertices")
                nodes = n
                edges += m
                break
    directed = 2 * edges if undirected else edges
    if not directed:
        raise ValueError("no graph found")
    rows, cols = [], []
    pending = [None] * len(texts)
    while True:
        smallest = -1
        for f, ls in enumerate(lines):
            while pending[f] is None and pos[f] < len(ls):
                line = ls[pos[f]]
                pos[f] += 1
                if line[0] != "a":
                    continue
                v = _scan_ints(line[1:], 2)
                if len(v) != 2:
                    raise ValueError("badly formed edge")
                u, w = v
                _check_id(u, nodes)
                _check_id(w, nodes)
                if len(rows) >= directed:
                    raise ValueError("more edges than declared")
                pending[f] = (u - 1, w - 1)
                if undirected:
                    rows.append(w - 1)
                    cols.append(u - 1)
            if pending[f] is not None and (smallest < 0 or pending[f][0] < pending[smallest][0]):
                smallest = f
        if smallest < 0:
            break
        rows.append(pending[smallest][0])
        cols.append(pending[smallest][1])
        pending[smallest] = None
    if len(rows) != directed:
        raise ValueError("only %d/%d edges read" % (len(rows), directed))
    return nodes, rows, cols
