// Graph files: MatrixMarket, METIS and DIMACS text parsed on the host into the
// edge order the reference's readers produce, then turned into a device CSR
// by the GPU stable sort by row (b40c_graph_from_coo) or, for METIS, which is
// already CSR, by a plain upload.  Host-only code: the parsers run without a
// GPU.
//
//  * MatrixMarket: ReadMarketStream / BuildMarketGraph
//    (b40c/graph/builder/market.cuh:50-158, 172-205)
//  * METIS: ReadMetisStream / BuildMetisGraph / WriteMetisStream
//    (b40c/graph/builder/metis.cuh:52-158, 184-212, 220-245)
//  * DIMACS: ReadDimacsStream / BuildDimacsGraph
//    (b40c/graph/builder/dimacs.cuh:54-220, 230-283)
//
// Parsing is line-parallel: the text is read whole, cut into chunks at line
// boundaries, each chunk parsed by its own thread, and the per-chunk results
// concatenated in file order, so the output does not depend on the number of
// threads.  Line handling follows the reference's fscanf("%[^\n]\n") loop:
// leading blanks of a line and blank lines are skipped.  Where the reference
// would index out of bounds or build a corrupt CSR (a vertex id of 0 or above
// the vertex count, more METIS vertex lines than vertices, an unsorted
// directed MatrixMarket file) these readers report an error or sort instead.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"

struct b40c_edge_list {
    int32_t nodes = 0;
    int32_t is_csr = 0;      // 1: METIS (a = offsets[nodes + 1], b = columns)
    std::vector<int32_t> a;  // COO rows in reference order, or CSR offsets
    std::vector<int32_t> b;  // COO columns, or CSR column indices
};

namespace b40c {
namespace {

constexpr int kMaxParseThreads = 16;

b40c_status load_text(const char* path, std::string& out) {
    FILE* f = path ? std::fopen(path, "rb") : stdin;
    if (!f) return fail(B40C_ERR_INVALID_ARG, "unable to open %s", path);
    char chunk[1 << 16];
    size_t got;
    out.clear();
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) out.append(chunk, got);
    const bool err = std::ferror(f) != 0;
    if (path) std::fclose(f);
    if (err) return fail(B40C_ERR_INVALID_ARG, "read error on %s", path ? path : "stdin");
    return B40C_OK;
}

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// scanf("%lld")-style: optional blanks, optional sign, decimal digits.
// Values beyond 19 digits saturate (they fail every range check anyway).
inline bool parse_ll(const char*& p, const char* e, long long& v) {
    while (p < e && is_blank(*p)) ++p;
    bool neg = false;
    if (p < e && (*p == '+' || *p == '-')) {
        neg = *p == '-';
        ++p;
    }
    if (p >= e || *p < '0' || *p > '9') return false;
    unsigned long long x = 0;
    int digits = 0;
    for (; p < e && *p >= '0' && *p <= '9'; ++p, ++digits)
        x = digits < 18 ? x * 10 + (unsigned long long)(*p - '0') : 0x7FFFFFFFFFFFFFFFull;
    v = neg ? -(long long)x : (long long)x;
    return true;
}

// Line [b, e) without its '\n'; `next` is where the following line starts.
inline void line_at(const char* p, const char* end, const char*& b, const char*& e, const char*& next) {
    b = p;
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
    e = nl ? nl : end;
    next = nl ? nl + 1 : end;
}

// Cuts [p, end) into at most `parts` chunks starting at line starts.
std::vector<const char*> split_lines(const char* p, const char* end, int parts) {
    std::vector<const char*> cuts{p};
    const size_t len = (size_t)(end - p);
    for (int i = 1; i < parts; ++i) {
        const char* c = p + len * (size_t)i / (size_t)parts;
        if (c <= cuts.back()) continue;
        const char* nl = static_cast<const char*>(std::memchr(c, '\n', (size_t)(end - c)));
        if (!nl) break;
        if (nl + 1 > cuts.back() && nl + 1 < end) cuts.push_back(nl + 1);
    }
    cuts.push_back(end);
    return cuts;
}

int parse_threads(size_t bytes) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t by_size = std::max<size_t>(1, bytes >> 20);  // about a MiB of text per thread at least
    return (int)std::min<size_t>({(size_t)kMaxParseThreads, (size_t)hw, by_size});
}

// Runs fn(i) for every chunk i, one thread per chunk.
template <typename Fn>
void run_chunks(int n, Fn fn) {
    if (n == 1) {
        fn(0);
        return;
    }
    std::vector<std::thread> th;
    for (int i = 0; i < n; ++i) th.emplace_back(fn, i);
    for (auto& t : th) t.join();
}

// Per-chunk parse result: vertex-id pairs in line order, or the first error.
struct PairChunk {
    std::vector<int32_t> u, v;
    long long err_line = -1;  // chunk-local line number of the first error
    std::string err;
};

// Data lines of a MatrixMarket body or DIMACS file: "x y ..." pairs (after
// the tag for DIMACS), each id in [1, n].  Other lines are skipped.
void parse_pairs(const char* p, const char* end, char tag, long long n, PairChunk& out) {
    long long line = 0;
    while (p < end) {
        const char *b, *e, *next;
        line_at(p, end, b, e, next);
        p = next;
        ++line;
        while (b < e && is_blank(*b)) ++b;
        if (b == e) continue;
        if (tag) {
            if (*b != tag) continue;  // comments and other descriptors
            ++b;
        } else if (*b == '%') {
            continue;
        }
        long long x, y;
        const char* q = b;
        if (!parse_ll(q, e, x) || !parse_ll(q, e, y)) {
            out.err_line = line;
            out.err = "badly formed edge";
            return;
        }
        if (x < 1 || x > n || y < 1 || y > n) {
            out.err_line = line;
            out.err = "vertex id out of range [1, " + std::to_string(n) + "]";
            return;
        }
        out.u.push_back((int32_t)(x - 1));
        out.v.push_back((int32_t)(y - 1));
    }
}

// Parses [p, end) in line-aligned chunks; returns the concatenated pairs or
// the first error (in file order).
b40c_status parse_pairs_parallel(const char* what, const char* p, const char* end, char tag, long long n,
                                 std::vector<int32_t>& u, std::vector<int32_t>& v) {
    const std::vector<const char*> cuts = split_lines(p, end, parse_threads((size_t)(end - p)));
    const int parts = (int)cuts.size() - 1;
    std::vector<PairChunk> res(parts);
    run_chunks(parts, [&](int i) { parse_pairs(cuts[i], cuts[i + 1], tag, n, res[i]); });
    size_t total = 0;
    for (const auto& r : res) {
        if (r.err_line >= 0) return fail(B40C_ERR_INVALID_ARG, "%s: %s", what, r.err.c_str());
        total += r.u.size();
    }
    u.resize(total);
    v.resize(total);
    std::vector<size_t> at(parts + 1, 0);
    for (int i = 0; i < parts; ++i) at[i + 1] = at[i] + res[i].u.size();
    run_chunks(parts, [&](int i) {
        std::copy(res[i].u.begin(), res[i].u.end(), u.begin() + at[i]);
        std::copy(res[i].v.begin(), res[i].v.end(), v.begin() + at[i]);
    });
    return B40C_OK;
}

// First non-blank, non-comment line at or after p (comment = `comment` char
// first); returns its bounds and the position after it, or false.
bool first_data_line(const char* p, const char* end, char comment, const char*& b, const char*& e,
                     const char*& next) {
    while (p < end) {
        line_at(p, end, b, e, next);
        p = next;
        while (b < e && is_blank(*b)) ++b;
        if (b < e && *b != comment) return true;
    }
    return false;
}

// ---- MatrixMarket (market.cuh:50-158) -----------------------------------
// Header "rows cols entries" (square); entry "i j [value]" is the edge
// (j-1) -> (i-1); undirected appends the reverse edge right after it.
b40c_status read_market(const std::string& text, bool undirected, b40c_edge_list& el) {
    const char* p = text.data();
    const char* end = p + text.size();
    const char *b, *e, *body;
    if (!first_data_line(p, end, '%', b, e, body)) return fail(B40C_ERR_INVALID_ARG, "MatrixMarket: no graph found");
    long long rows, cols, nnz;
    const char* q = b;
    if (!parse_ll(q, e, rows) || !parse_ll(q, e, cols) || !parse_ll(q, e, nnz))
        return fail(B40C_ERR_INVALID_ARG, "MatrixMarket: invalid problem description");
    if (rows != cols)
        return fail(B40C_ERR_INVALID_ARG, "MatrixMarket: not square (%lld, %lld)", rows, cols);
    const long long directed = undirected ? 2 * nnz : nnz;
    if (rows < 1 || rows > 0x7FFFFFFFll || nnz < 0 || directed > 0x7FFFFFFFll)
        return fail(B40C_ERR_INVALID_ARG, "MatrixMarket: sizes (%lld vertices, %lld entries) outside int32", rows, nnz);
    std::vector<int32_t> x, y;
    B40C_TRY(parse_pairs_parallel("MatrixMarket", body, end, 0, rows, x, y));
    if ((long long)x.size() != nnz)
        return fail(B40C_ERR_INVALID_ARG, "MatrixMarket: %lld/%lld edges read", (long long)x.size(), nnz);
    el.nodes = (int32_t)rows;
    el.is_csr = 0;
    el.a.resize((size_t)directed);
    el.b.resize((size_t)directed);
    const size_t step = undirected ? 2 : 1;
    for (size_t i = 0; i < x.size(); ++i) {
        el.a[i * step] = y[i];
        el.b[i * step] = x[i];
        if (undirected) {
            el.a[i * step + 1] = x[i];
            el.b[i * step + 1] = y[i];
        }
    }
    return B40C_OK;
}

// ---- METIS (metis.cuh:52-158) --------------------------------------------
// Header "n m [fmt ...]"; then line k lists vertex k's neighbours (1-based).
// A '%' comment runs to the end of its line and takes that line's newline
// with it, so a comment line is not a vertex line.  The file may list every
// undirected edge twice (2m entries) or give the directed count (m entries).
struct MetisChunk {
    std::vector<int32_t> cols;
    std::vector<int64_t> bounds;  // local column count at each vertex-line end
    long long err_line = -1;
    std::string err;
};

void parse_metis_chunk(const char* p, const char* end, long long n, MetisChunk& out) {
    long long line = 0;
    while (p < end) {
        const char *b, *e, *next;
        line_at(p, end, b, e, next);
        p = next;
        ++line;
        bool comment = false;
        const char* q = b;
        while (true) {
            while (q < e && is_blank(*q)) ++q;
            if (q >= e) break;
            if (*q == '%') {
                comment = true;
                break;
            }
            long long x;
            if (!parse_ll(q, e, x)) {
                out.err_line = line;
                out.err = "invalid format";
                return;
            }
            if (x < 1 || x > n) {
                out.err_line = line;
                out.err = "edge to " + std::to_string(x) + " outside [1, " + std::to_string(n) + "]";
                return;
            }
            out.cols.push_back((int32_t)(x - 1));
        }
        // The newline ends a vertex line unless a comment consumed it (the
        // last line of a file without a final newline has none).
        if (!comment && e < end) out.bounds.push_back((int64_t)out.cols.size());
    }
}

b40c_status read_metis(const std::string& text, b40c_edge_list& el) {
    const char* p = text.data();
    const char* end = p + text.size();
    const char *b, *e, *body;
    if (!first_data_line(p, end, '%', b, e, body)) return fail(B40C_ERR_INVALID_ARG, "METIS: no graph found");
    long long n, m, fmt = 0;
    const char* q = b;
    if (!parse_ll(q, e, n) || !parse_ll(q, e, m)) return fail(B40C_ERR_INVALID_ARG, "METIS: invalid format");
    if (parse_ll(q, e, fmt) && fmt != 0)
        return fail(B40C_ERR_UNSUPPORTED, "METIS: weighted format %lld (only unweighted graphs are read)", fmt);
    if (n < 1 || n > 0x7FFFFFFFll || m < 0 || 2 * m > 0x7FFFFFFFll)
        return fail(B40C_ERR_INVALID_ARG, "METIS: sizes (%lld vertices, %lld edges) outside int32", n, m);
    // The header line's own newline starts vertex 0 (offsets[0] = 0); a file
    // that ends with the header has only isolated vertices.
    const std::vector<const char*> cuts = split_lines(body, end, parse_threads((size_t)(end - body)));
    const int parts = (int)cuts.size() - 1;
    std::vector<MetisChunk> res(parts);
    run_chunks(parts, [&](int i) { parse_metis_chunk(cuts[i], cuts[i + 1], n, res[i]); });
    size_t total = 0;
    for (const auto& r : res) {
        if (r.err_line >= 0) return fail(B40C_ERR_INVALID_ARG, "METIS: %s", r.err.c_str());
        total += r.cols.size();
    }
    if ((long long)total != 2 * m && (long long)total != m)
        return fail(B40C_ERR_INVALID_ARG, "METIS: %zu edges read, header names %lld undirected (%lld or %lld entries)",
                    total, m, m, 2 * m);
    el.nodes = (int32_t)n;
    el.is_csr = 1;
    el.b.resize(total);
    el.a.assign((size_t)n + 1, (int32_t)total);
    el.a[0] = 0;
    // Vertex line k (k >= 1 boundaries seen) starts at the k-th boundary.
    size_t k = 0, base = 0;
    for (const auto& r : res) {
        std::copy(r.cols.begin(), r.cols.end(), el.b.begin() + base);
        for (int64_t c : r.bounds) {
            ++k;
            if (k <= (size_t)n) {
                el.a[k] = (int32_t)(base + (size_t)c);
            } else if (base + (size_t)c != total) {
                return fail(B40C_ERR_INVALID_ARG, "METIS: more vertex lines than the %lld vertices", n);
            }
        }
        base += r.cols.size();
    }
    // Lines past the last vertex may be blank, but carry no edges.
    if (k >= (size_t)n && el.a[(size_t)n] != (int32_t)total)
        return fail(B40C_ERR_INVALID_ARG, "METIS: more vertex lines than the %lld vertices", n);
    return B40C_OK;
}

// ---- DIMACS (dimacs.cuh:54-220) ------------------------------------------
// Per file: lines before the "p <type> n m" line are skipped; "a u v [w]"
// lines after it are edges.  Files of a splice are merged by row: each step
// takes the pending edge with the smallest row (ties: lowest file), and in
// undirected mode an edge's reverse is emitted as soon as its line is read.
b40c_status read_dimacs(const std::vector<std::string>& texts, bool undirected, b40c_edge_list& el) {
    long long nodes = 0, edges = 0;
    std::vector<std::vector<int32_t>> fu(texts.size()), fv(texts.size());
    for (size_t f = 0; f < texts.size(); ++f) {
        const char* p = texts[f].data();
        const char* end = p + texts[f].size();
        const char *b = nullptr, *e = nullptr, *body = end;
        bool found = false;
        while (p < end) {
            line_at(p, end, b, e, body);
            p = body;
            while (b < e && is_blank(*b)) ++b;
            if (b < e && *b == 'p') {
                found = true;
                break;
            }
        }
        if (!found) continue;  // no problem line: the file adds nothing
        const char* q = b + 1;
        while (q < e && is_blank(*q)) ++q;
        while (q < e && !is_blank(*q)) ++q;  // problem type
        long long n, m;
        if (!parse_ll(q, e, n) || !parse_ll(q, e, m) || n < 1 || m < 0)
            return fail(B40C_ERR_INVALID_ARG, "DIMACS: invalid problem line");
        if (nodes && nodes != n)
            return fail(B40C_ERR_INVALID_ARG, "DIMACS: splice files do not name the same number of vertices");
        nodes = n;
        edges += m;
        B40C_TRY(parse_pairs_parallel("DIMACS", body, end, 'a', n, fu[f], fv[f]));
    }
    const long long directed = undirected ? 2 * edges : edges;
    if (!nodes || !directed) return fail(B40C_ERR_INVALID_ARG, "DIMACS: no graph found");
    if (nodes > 0x7FFFFFFFll || directed > 0x7FFFFFFFll)
        return fail(B40C_ERR_INVALID_ARG, "DIMACS: sizes (%lld vertices, %lld edges) outside int32", nodes, directed);
    size_t read = 0;
    for (size_t f = 0; f < texts.size(); ++f) read += fu[f].size();
    if ((long long)read * (undirected ? 2 : 1) != directed)
        return fail(B40C_ERR_INVALID_ARG, "DIMACS: %lld/%lld edges read", (long long)read * (undirected ? 2 : 1),
                    directed);
    el.nodes = (int32_t)nodes;
    el.is_csr = 0;
    el.a.clear();
    el.b.clear();
    el.a.reserve((size_t)directed);
    el.b.reserve((size_t)directed);
    const size_t F = texts.size();
    std::vector<size_t> pos(F, 0);
    std::vector<long long> pend(F, -1);
    while (true) {
        int smallest = -1;
        for (size_t i = 0; i < F; ++i) {
            if (pend[i] < 0 && pos[i] < fu[i].size()) {
                pend[i] = (long long)pos[i]++;
                if (undirected) {
                    el.a.push_back(fv[i][pend[i]]);
                    el.b.push_back(fu[i][pend[i]]);
                }
            }
            if (pend[i] >= 0 && (smallest < 0 || fu[i][pend[i]] < fu[smallest][pend[smallest]])) smallest = (int)i;
        }
        if (smallest < 0) break;
        el.a.push_back(fu[smallest][pend[smallest]]);
        el.b.push_back(fv[smallest][pend[smallest]]);
        pend[smallest] = -1;
    }
    return B40C_OK;
}

}  // namespace
}  // namespace b40c

using b40c::fail;

extern "C" {

b40c_status b40c_edge_list_read(const char* path, int32_t format, int32_t undirected, int32_t splice,
                                b40c_edge_list** out) {
    if (!out) return fail(B40C_ERR_INVALID_ARG, "b40c_edge_list_read: null out");
    if (format < B40C_FILE_MARKET || format > B40C_FILE_DIMACS)
        return fail(B40C_ERR_INVALID_ARG, "b40c_edge_list_read: unknown format %d", format);
    if (splice < 0 || (splice > 0 && (format != B40C_FILE_DIMACS || !path)))
        return fail(B40C_ERR_INVALID_ARG, "b40c_edge_list_read: splice applies to DIMACS files only");
    b40c_edge_list* el = new b40c_edge_list();
    b40c_status s = B40C_OK;
    if (format == B40C_FILE_DIMACS) {
        std::vector<std::string> texts;
        if (splice > 0) {
            // path.0, path.1, ... up to the first that does not open.
            for (int i = 0; i < splice; ++i) {
                const std::string name = std::string(path) + "." + std::to_string(i);
                FILE* f = std::fopen(name.c_str(), "rb");
                if (!f) break;
                std::fclose(f);
                texts.emplace_back();
                s = b40c::load_text(name.c_str(), texts.back());
                if (s != B40C_OK) break;
            }
            if (s == B40C_OK && texts.empty()) s = fail(B40C_ERR_INVALID_ARG, "DIMACS: no splice file %s.0", path);
        } else {
            texts.emplace_back();
            s = b40c::load_text(path, texts.back());
        }
        if (s == B40C_OK) s = b40c::read_dimacs(texts, undirected != 0, *el);
    } else {
        std::string text;
        s = b40c::load_text(path, text);
        if (s == B40C_OK)
            s = (format == B40C_FILE_MARKET) ? b40c::read_market(text, undirected != 0, *el) : b40c::read_metis(text, *el);
    }
    if (s != B40C_OK) {
        delete el;
        return s;
    }
    *out = el;
    return B40C_OK;
}

b40c_status b40c_edge_list_info(const b40c_edge_list* el, int32_t* nodes, int64_t* edges, int32_t* is_csr) {
    if (!el) return fail(B40C_ERR_INVALID_ARG, "b40c_edge_list_info: null");
    if (nodes) *nodes = el->nodes;
    if (edges) *edges = (int64_t)el->b.size();
    if (is_csr) *is_csr = el->is_csr;
    return B40C_OK;
}

b40c_status b40c_edge_list_copy(const b40c_edge_list* el, int32_t* rows_or_offsets, int32_t* cols) {
    if (!el) return fail(B40C_ERR_INVALID_ARG, "b40c_edge_list_copy: null");
    if (rows_or_offsets && !el->a.empty()) std::memcpy(rows_or_offsets, el->a.data(), el->a.size() * 4);
    if (cols && !el->b.empty()) std::memcpy(cols, el->b.data(), el->b.size() * 4);
    return B40C_OK;
}

void b40c_edge_list_destroy(b40c_edge_list* el) { delete el; }

b40c_status b40c_graph_from_edge_list(const b40c_edge_list* el, b40c_graph** out) {
    if (!el || !out) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_from_edge_list: null argument");
    const int64_t m = (int64_t)el->b.size();
    if (el->is_csr) return b40c_graph_from_host_csr(el->nodes, m, el->a.data(), el->b.data(), out);
    return b40c_graph_from_coo(el->nodes, m, el->a.data(), el->b.data(), out);
}

b40c_status b40c_graph_read(const char* path, int32_t format, int32_t undirected, int32_t splice, b40c_graph** out) {
    b40c_edge_list* el = nullptr;
    B40C_TRY(b40c_edge_list_read(path, format, undirected, splice, &el));
    const b40c_status s = b40c_graph_from_edge_list(el, out);
    b40c_edge_list_destroy(el);
    return s;
}

b40c_status b40c_graph_write_metis(const b40c_graph* g, const char* path) {
    if (!g) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_write_metis: null graph");
    std::vector<int32_t> off((size_t)g->nodes + 1), col((size_t)g->edges);
    B40C_TRY(b40c_graph_copy_to_host(g, off.data(), col.data()));
    FILE* f = path ? std::fopen(path, "wb") : stdout;
    if (!f) return fail(B40C_ERR_INVALID_ARG, "unable to open %s for writing", path);
    std::string buf;
    buf.reserve(1 << 20);
    buf += std::to_string(g->nodes) + " " + std::to_string(g->edges) + "\n";
    bool ok = true;
    for (int32_t v = 0; v < g->nodes && ok; ++v) {
        for (int32_t e = off[v]; e < off[v + 1]; ++e) {
            buf += std::to_string((long long)col[e] + 1);
            buf += ' ';
        }
        buf += '\n';
        if (buf.size() > (1u << 20)) {
            ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
            buf.clear();
        }
    }
    if (ok && !buf.empty()) ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (path) ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(B40C_ERR_INVALID_ARG, "write error on %s", path ? path : "stdout");
    return B40C_OK;
}

}  // extern "C"
