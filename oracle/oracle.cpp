// CPU oracle for the single-GPU BFS path — TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker (or the timed CPU baseline).
// The product (back40computing_amd) never links, imports or calls it.
//
// Every function restates the reference algorithm it follows (file:line in
// the b40c tree) in plain serial C++, so that the GPU path can be compared
// with it bit for bit.  Each one is pinned by checks in tests/ that do not
// retype its formula (closed forms, brute force on tiny graphs, library
// routines, invariants); see the per-function notes.
//
// Generator definitions (RMAT edge stream, lattice permutation, source
// choice) are the synthetic-input definitions of this project.  They are
// counter-based (SplitMix64, Steele/Lea/Flood 2014) so the GPU builder can
// produce the identical stream in parallel.

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>
#include <sys/resource.h>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint32_t prob_threshold(double p) {
    double t = p * 4294967296.0;
    if (!(t > 0.0)) return 0u;
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)t;
}

inline int ceil_log2_u64(uint64_t n) {
    int b = 0;
    while ((1ull << b) < n) ++b;
    return b;
}

// Serial CSR fill from rows sorted ascending; restates the offset loop of
// CsrGraph::FromCoo (b40c/graph/csr_graph.cuh:127-147): rows up to and
// including the current row get the current edge index, trailing rows and
// the end-of-list element get `edges`.
void fill_offsets(int32_t n, int64_t m, const int32_t* sorted_rows, int32_t* offsets) {
    int64_t prev_row = -1;
    for (int64_t e = 0; e < m; ++e) {
        int64_t r = sorted_rows[e];
        for (int64_t row = prev_row + 1; row <= r; ++row) offsets[row] = (int32_t)e;
        prev_row = r;
    }
    for (int64_t row = prev_row + 1; row <= n; ++row) offsets[row] = (int32_t)m;
}

// Unique (src,dst) keys, packed (src << shift) | dst and sorted, to CSR.
int64_t keys_to_csr(int32_t n, int shift, std::vector<uint64_t>& keys, int32_t* offsets, int32_t* cols) {
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const uint64_t mask = (shift >= 64) ? ~0ull : ((1ull << shift) - 1);
    int64_t m = (int64_t)keys.size();
    std::vector<int32_t> rows(m);
    for (int64_t e = 0; e < m; ++e) {
        rows[e] = (int32_t)(keys[e] >> shift);
        cols[e] = (int32_t)(keys[e] & mask);
    }
    fill_offsets(n, m, rows.data(), offsets);
    return m;
}

}  // namespace

extern "C" {

uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }

// Quadrant thresholds of the RMAT recursion as 32-bit integers:
// t[0] = a, t[1] = a + b, t[2] = a + b + c, each scaled by 2^32 and truncated.
void orc_rmat_thresholds(double a, double b, double c, uint32_t* t) {
    t[0] = prob_threshold(a);
    t[1] = prob_threshold(a + b);
    t[2] = prob_threshold(a + b + c);
}

// Raw sampled RMAT edges [i0, i1) of the stream (no symmetrization, self
// loops kept).  Quadrant semantics follow ChoosePartition
// (b40c/graph/builder/rmat.cuh:55-81): a -> no change, b -> column += step,
// c -> row += step, d -> both; ids are 0-based and no parameter noise is
// applied (the configs define a noise-free generator).  Draw l of edge i is
// 32-bit half (l & 1) of splitmix64(splitmix64(seed) + 16*i + l/2).
void orc_rmat_edges(int scale, int64_t i0, int64_t i1, const uint32_t* t, uint64_t seed,
                    int32_t* u_out, int32_t* v_out) {
    const uint64_t smix = splitmix64(seed);
    for (int64_t i = i0; i < i1; ++i) {
        uint32_t u = 0, v = 0;
        uint64_t x = 0;
        for (int l = 0; l < scale; ++l) {
            if ((l & 1) == 0) x = splitmix64(smix + 16ull * (uint64_t)i + (uint64_t)(l >> 1));
            uint32_t r = (l & 1) ? (uint32_t)(x >> 32) : (uint32_t)x;
            uint32_t step = 1u << (scale - 1 - l);
            if (r < t[0]) {
            } else if (r < t[1]) {
                v += step;
            } else if (r < t[2]) {
                u += step;
            } else {
                u += step;
                v += step;
            }
        }
        u_out[i - i0] = (int32_t)u;
        v_out[i - i0] = (int32_t)v;
    }
}

// Symmetrized, self-loop-free, deduplicated RMAT CSR (int32).  Sorting the
// unique packed (src,dst) keys with std::sort is the CSR-order oracle
// (test/radix_sort/test_sort.cu:254-255 checks the GPU sort the same way).
// `cols` must hold 2 * edge_factor * 2^scale entries.  Returns E.
int64_t orc_build_rmat_csr(int scale, int edge_factor, double a, double b, double c,
                           uint64_t seed, int32_t* offsets, int32_t* cols) {
    const int32_t n = 1 << scale;
    const int64_t m = (int64_t)edge_factor << scale;
    uint32_t t[3];
    orc_rmat_thresholds(a, b, c, t);
    std::vector<int32_t> u(m), v(m);
    orc_rmat_edges(scale, 0, m, t, seed, u.data(), v.data());
    std::vector<uint64_t> keys;
    keys.reserve(2 * m);
    for (int64_t i = 0; i < m; ++i) {
        if (u[i] == v[i]) continue;
        keys.push_back(((uint64_t)u[i] << scale) | (uint64_t)v[i]);
        keys.push_back(((uint64_t)v[i] << scale) | (uint64_t)u[i]);
    }
    return keys_to_csr(n, scale, keys, offsets, cols);
}

// Random relabelling of the w^3 lattice: original id o = (i*w + j)*w + k
// (BuildGrid3dGraph, b40c/graph/builder/grid3d.cuh:73-77) gets new id
// perm[o] = rank of key(o) = (splitmix64(splitmix64(seed) + o) with the low
// `bits` cleared) | o, bits = ceil(log2(w^3)).
void orc_lattice_perm(int w, uint64_t seed, int32_t* perm) {
    const int64_t n = (int64_t)w * w * w;
    const int bits = ceil_log2_u64((uint64_t)n);
    const uint64_t lowmask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    const uint64_t smix = splitmix64(seed);
    std::vector<uint64_t> keys(n);
    for (int64_t o = 0; o < n; ++o) keys[o] = (splitmix64(smix + (uint64_t)o) & ~lowmask) | (uint64_t)o;
    std::sort(keys.begin(), keys.end());
    for (int64_t r = 0; r < n; ++r) perm[keys[r] & lowmask] = (int32_t)r;
}

// 6-neighbour w^3 lattice in the reference's neighbour order (k-1, k+1,
// j-1, j+1, i-1, i+1; grid3d.cuh:79-113) without the per-vertex self loop
// (grid3d.cuh:115-117 is dropped by the high-diameter config), relabelled by
// orc_lattice_perm and stored as CSR with ascending columns.  `perm` (w^3)
// receives the relabelling.  Returns E = 6*w^2*(w-1).
int64_t orc_build_lattice_csr(int w, uint64_t seed, int32_t* offsets, int32_t* cols, int32_t* perm) {
    const int64_t n = (int64_t)w * w * w;
    orc_lattice_perm(w, seed, perm);
    const int shift = ceil_log2_u64((uint64_t)n);
    std::vector<uint64_t> keys;
    keys.reserve(6 * n);
    for (int64_t i = 0; i < w; ++i)
        for (int64_t j = 0; j < w; ++j)
            for (int64_t k = 0; k < w; ++k) {
                const int64_t me = (i * w + j) * w + k;
                const uint64_t p = (uint64_t)perm[me];
                auto add = [&](int64_t o) { keys.push_back((p << shift) | (uint64_t)perm[o]); };
                if (k - 1 >= 0) add(me - 1);
                if (k + 1 < w) add(me + 1);
                if (j - 1 >= 0) add(me - w);
                if (j + 1 < w) add(me + w);
                if (i - 1 >= 0) add(me - (int64_t)w * w);
                if (i + 1 < w) add(me + (int64_t)w * w);
            }
    return keys_to_csr((int32_t)n, shift, keys, offsets, cols);
}

// Closed-form BFS depth on the relabelled lattice: Manhattan distance of the
// unpermuted coordinates.  `inv` maps new id -> original id.
void orc_lattice_depth(int w, const int32_t* inv, int32_t src, int32_t* depth) {
    const int64_t n = (int64_t)w * w * w;
    const int64_t os = inv[src];
    const int64_t si = os / ((int64_t)w * w), sj = (os / w) % w, sk = os % w;
    for (int64_t v = 0; v < n; ++v) {
        const int64_t o = inv[v];
        const int64_t i = o / ((int64_t)w * w), j = (o / w) % w, k = o % w;
        depth[v] = (int32_t)(std::llabs(i - si) + std::llabs(j - sj) + std::llabs(k - sk));
    }
}

// CSR from an unordered COO edge list with the reference's semantics:
// std::stable_sort by row only, columns kept in input order within a row
// (CsrGraph::FromCoo + DimacsTupleCompare, csr_graph.cuh:108-152,
// coo_edge_tuple.cuh:82-98).  Returns m.
int64_t orc_csr_from_coo(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols,
                         int32_t* offsets, int32_t* out_cols) {
    std::vector<int64_t> idx(m);
    for (int64_t e = 0; e < m; ++e) idx[e] = e;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return rows[x] < rows[y]; });
    std::vector<int32_t> srows(m);
    for (int64_t e = 0; e < m; ++e) {
        srows[e] = rows[idx[e]];
        out_cols[e] = cols[idx[e]];
    }
    fill_offsets(n, m, srows.data(), offsets);
    return m;
}

// Serial FIFO BFS, restating SimpleReferenceBfs (test/bfs/test_bfs.cu:507-578):
// labels := -1, label[src] := 0; each dequeued vertex labels its unlabelled
// neighbours with its label + 1 and enqueues them.  `pred` (nullable) gets the
// discovering vertex (the source gets -2, the reference's source predecessor,
// contract_expand_atomic/kernel.cuh:321-325).  Returns search depth =
// max label + 1 (test_bfs.cu:550-559).
int32_t orc_bfs(int32_t n, const int32_t* offsets, const int32_t* cols, int32_t src,
                int32_t* depth, int32_t* pred) {
    for (int32_t i = 0; i < n; ++i) depth[i] = -1;
    if (pred)
        for (int32_t i = 0; i < n; ++i) pred[i] = -1;
    depth[src] = 0;
    if (pred) pred[src] = -2;
    int32_t search_depth = 0;
    std::deque<int32_t> frontier;
    frontier.push_back(src);
    while (!frontier.empty()) {
        const int32_t u = frontier.front();
        frontier.pop_front();
        const int32_t nd = depth[u] + 1;
        for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
            const int32_t v = cols[e];
            if (depth[v] == -1) {
                depth[v] = nd;
                if (pred) pred[v] = u;
                if (search_depth < nd) search_depth = nd;
                frontier.push_back(v);
            }
        }
    }
    return search_depth + 1;
}

// Same BFS timed as the reference times its CPU path: wall clock
// (steady_clock) and user CPU time from getrusage (CpuTimer,
// test/b40c_test_util.h:503-522).  Initialisation is outside the timer.
int32_t orc_bfs_timed(int32_t n, const int32_t* offsets, const int32_t* cols, int32_t src,
                      int32_t* depth, double* wall_ms, double* user_ms) {
    for (int32_t i = 0; i < n; ++i) depth[i] = -1;
    depth[src] = 0;
    int32_t search_depth = 0;
    std::deque<int32_t> frontier;
    rusage r0, r1;
    getrusage(RUSAGE_SELF, &r0);
    auto t0 = std::chrono::steady_clock::now();
    frontier.push_back(src);
    while (!frontier.empty()) {
        const int32_t u = frontier.front();
        frontier.pop_front();
        const int32_t nd = depth[u] + 1;
        for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
            const int32_t v = cols[e];
            if (depth[v] == -1) {
                depth[v] = nd;
                if (search_depth < nd) search_depth = nd;
                frontier.push_back(v);
            }
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    getrusage(RUSAGE_SELF, &r1);
    *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    *user_ms = (r1.ru_utime.tv_sec - r0.ru_utime.tv_sec) * 1e3 + (r1.ru_utime.tv_usec - r0.ru_utime.tv_usec) * 1e-3;
    return search_depth + 1;
}

// Depth labels for many sources, one serial BFS per source, spread over
// `threads` host threads (validation throughput only; each BFS is serial).
void orc_bfs_many(int32_t n, const int32_t* offsets, const int32_t* cols, const int32_t* srcs,
                  int32_t nsrc, int32_t* depths, int32_t threads) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int32_t t = 0; t < threads; ++t) {
        pool.emplace_back([=]() {
            for (int32_t q = t; q < nsrc; q += threads)
                orc_bfs(n, offsets, cols, srcs[q], depths + (int64_t)q * n, nullptr);
        });
    }
    for (auto& th : pool) th.join();
}

// Predecessor validity, restating DisplayStats' --mark-pred checks
// (test/bfs/test_bfs.cu:387-438): a parentless node must have reference
// depth <= 0; a parent must have depth one less than the node; the parent
// must list the node in its adjacency.  Additionally the source must carry
// -2 and an unreached node -1 (label conventions of the path).  Returns -1
// when valid, else the first offending node.
int64_t orc_check_pred(int32_t n, const int32_t* offsets, const int32_t* cols,
                       const int32_t* ref_depth, const int32_t* pred) {
    for (int32_t node = 0; node < n; ++node) {
        const int32_t parent = pred[node];
        if (parent < 0) {
            if (ref_depth[node] > 0) return node;
            if (ref_depth[node] == 0 && parent != -2) return node;
            if (ref_depth[node] < 0 && parent != -1) return node;
            continue;
        }
        if (parent >= n) return node;
        if (ref_depth[parent] + 1 != ref_depth[node]) return node;
        bool found = false;
        for (int32_t e = offsets[parent]; e < offsets[parent + 1]; ++e)
            if (cols[e] == node) {
                found = true;
                break;
            }
        if (!found) return node;
    }
    return -1;
}

// edges_visited = sum of degrees of labelled vertices, nodes_visited = count
// (DisplayStats, test/bfs/test_bfs.cu:359-367).
int64_t orc_edges_visited(int32_t n, const int32_t* offsets, const int32_t* labels, int64_t* nodes_visited) {
    int64_t e = 0, v = 0;
    for (int32_t i = 0; i < n; ++i)
        if (labels[i] > -1) {
            ++v;
            e += offsets[i + 1] - offsets[i];
        }
    if (nodes_visited) *nodes_visited = v;
    return e;
}

// Work histogram by depth (Histogram, test_bfs.cu:246-337), from a
// search's depth labels: discovered[d] = vertices at depth d;
// expanded[d + 1] = degree sum of the depth-d vertices (the edges a
// level-synchronous search expands to reach depth d + 1); unique_expanded
// [d + 1] = distinct endpoints of those edges; level 0 expands the source
// alone (expanded = unique = 1).  Arrays hold search_depth + 1 entries.
// Returns 0, or -1 for a label outside [-1, search_depth).
int32_t orc_histogram(int32_t n, const int32_t* offsets, const int32_t* cols, const int32_t* labels,
                      int32_t search_depth, int64_t* expanded, int64_t* unique_expanded, int64_t* discovered) {
    const int32_t D = search_depth;
    for (int32_t d = 0; d <= D; ++d) expanded[d] = unique_expanded[d] = discovered[d] = 0;
    if (D < 1) return 0;
    expanded[0] = unique_expanded[0] = 1;
    std::vector<std::vector<int32_t>> frontier(D);
    for (int32_t v = 0; v < n; ++v) {
        const int32_t d = labels[v];
        if (d < -1 || d >= D) return -1;
        if (d < 0) continue;
        ++discovered[d];
        expanded[d + 1] += offsets[v + 1] - offsets[v];
        frontier[d].insert(frontier[d].end(), cols + offsets[v], cols + offsets[v + 1]);
    }
    for (int32_t d = 0; d < D; ++d) {
        std::sort(frontier[d].begin(), frontier[d].end());
        unique_expanded[d + 1] = std::unique(frontier[d].begin(), frontier[d].end()) - frontier[d].begin();
    }
    return 0;
}

// Seeded source choice among vertices of nonzero degree: candidate k is
// splitmix64(splitmix64(seed ^ 0x5EED) + k) mod n, accepted when its degree
// is positive and it was not chosen before (the reference re-rolls sources
// that reach fewer than 5 vertices, test_bfs.cu:444-445, 662-666; the
// configs ask for nonzero degree instead).  Returns the number chosen.
int32_t orc_pick_sources(int32_t n, const int32_t* offsets, uint64_t seed, int32_t count, int32_t* out) {
    const uint64_t smix = splitmix64(seed ^ 0x5EEDull);
    int32_t got = 0;
    std::vector<char> used(n, 0);
    for (uint64_t k = 0; got < count && k < 64ull * (uint64_t)n + 1024; ++k) {
        const int32_t c = (int32_t)(splitmix64(smix + k) % (uint64_t)n);
        if (offsets[c + 1] - offsets[c] > 0 && !used[c]) {
            used[c] = 1;
            out[got++] = c;
        }
    }
    return got;
}

// Radix-sort oracles: std::sort of u64 keys (test_sort.cu:254-255) and a
// stable sort of (u32 key, u32 value) pairs by key.
void orc_sort_u64(uint64_t* keys, int64_t n) { std::sort(keys, keys + n); }

void orc_stable_sort_pairs_u32(uint32_t* keys, uint32_t* vals, int64_t n) {
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return keys[x] < keys[y]; });
    std::vector<uint32_t> k2(n), v2(n);
    for (int64_t i = 0; i < n; ++i) {
        k2[i] = keys[idx[i]];
        v2[i] = vals[idx[i]];
    }
    std::memcpy(keys, k2.data(), n * sizeof(uint32_t));
    std::memcpy(vals, v2.data(), n * sizeof(uint32_t));
}

}  // extern "C"
