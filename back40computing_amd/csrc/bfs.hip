// Single-GPU BFS for gfx950: level kernels, persistent (cooperative) kernels,
// the CsrProblem object and the enactors.
//
// Reference counterparts (b40c):
//   CsrProblem                b40c/graph/bfs/csr_problem.cuh
//   EnactorExpandContract     b40c/graph/bfs/enactor_expand_contract.cuh
//   EnactorContractExpand     b40c/graph/bfs/enactor_contract_expand.cuh
//   EnactorTwoPhase           b40c/graph/bfs/enactor_two_phase.cuh
//   EnactorHybrid             b40c/graph/bfs/enactor_hybrid.cuh
//   level kernels             b40c/graph/bfs/{expand_contract_atomic,
//                             contract_expand_atomic, two_phase}/kernel.cuh
//
// Queue protocol: a "phase" is one kernel step that reads one frontier queue
// and writes the next.  Phase ph reads buffer ph&1 with its length in
// counter slot ph&3, appends to buffer (ph+1)&1 / slot (ph+1)&3 and clears
// slot (ph+2)&3.  Expand-contract and contract-expand use one phase per BFS
// level, two-phase uses two (contract, expand) or three (filter, contract,
// expand).  Depth labels are carried separately from the phase index.
#include <algorithm>
#include <cstring>
#include <vector>

#include "bfs_device.hpp"
#include "internal.hpp"

namespace b40c {

// ===========================================================================
// Level bodies (device).
// ===========================================================================

struct LevelShared {
    union {
        LdsQueue<false> qv;
        LdsQueue<true> qp;
    };
    GatherSmem g;
    HistoryHash hist;
    unsigned long long red[BFS_WAVES];
};

__device__ __forceinline__ void shared_init(LevelShared& s) {
    if (threadIdx.x == 0) {
        s.qv.count = 0;
        s.qp.count = 0;
    }
    gather_init(s.g);
    history_init(s.hist);
    __syncthreads();
}

__device__ __forceinline__ uint32_t phase_input_len(const BfsArgs& a, int ph) {
    uint32_t n = ld_relaxed(&a.ctr->qlen[ph & 3]);
    if (n > a.qcap[ph & 1]) n = 0;  // an overflowed queue counts as empty
    return n;
}

__device__ __forceinline__ void phase_begin(const BfsArgs& a, int ph, uint32_t n_in) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st_relaxed(&a.ctr->qlen[(ph + 2) & 3], 0u);
        a.ctr->level_edges[(ph + 2) & 3] = 0ull;
        if (n_in) atomic_add_u64(&a.ctr->queued, n_in);
    }
}

__device__ __forceinline__ GlobalQueue out_queue(const BfsArgs& a, int ph) {
    GlobalQueue q;
    q.keys = a.qk[(ph + 1) & 1];
    q.vals = a.qv[(ph + 1) & 1];
    q.len = &a.ctr->qlen[(ph + 1) & 3];
    q.cap = a.qcap[(ph + 1) & 1];
    q.overflow = &a.ctr->overflow;
    return q;
}

// Frontier entries are validated before they are used as addresses: an id
// outside [0, n) is dropped and flagged, so a corrupt queue surfaces as an
// error status instead of an out-of-bounds access.
__device__ __forceinline__ bool valid_vertex(const BfsArgs& a, int32_t v) {
    if ((uint32_t)v < (uint32_t)a.n) return true;
    st_relaxed(&a.ctr->corrupt, 1u);
    return false;
}

// Workgroup sum of a per-thread u64, added once to *dst.
__device__ __forceinline__ void wg_add_u64(unsigned long long v, unsigned long long* dst, LevelShared& s) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    __syncthreads();
    if (lane_id() == 0) s.red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < BFS_WAVES; ++w) t += s.red[w];
        if (t) atomic_add_u64(dst, t);
    }
}

// Expand-contract (vertex frontier -> vertex frontier).  Neighbours are
// culled and labelled at discovery (depth = depth_next), as in
// expand_contract_atomic/cta.cuh:212-214, and compacted into the next vertex
// frontier.  With COUNT_EDGES the degree sum of the discovered vertices is
// accumulated for the direction-optimizing switch.
template <bool COUNT_EDGES>
__device__ void td_phase(const BfsArgs& a, int ph, int32_t depth_next, LevelShared& s) {
    const uint32_t n_in = phase_input_len(a, ph);
    phase_begin(a, ph, n_in);
    const int32_t* qin = a.qk[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    bool found = false;
    unsigned long long eacc = 0;
    for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_in; base += gridDim.x * BFS_THREADS) {
        const uint32_t i = base + threadIdx.x;
        int32_t u = -1, start = 0, deg = 0;
        if (i < n_in) {
            u = qin[i];
            if (valid_vertex(a, u)) {
                start = a.offs[u];
                deg = a.offs[u + 1] - start;
            }
        }
        auto sink = [&](bool valid, int32_t v, int32_t src) {
            const bool won = valid && claim_exact(a.visited, v);
            if (won) {
                a.depth[v] = depth_next;
                a.pred[v] = src;
                found = true;
                if (COUNT_EDGES) eacc += (unsigned long long)(a.offs[v + 1] - a.offs[v]);
            }
            lq_append(s.qv, out, won, v, 0);
        };
        tiered_gather(a.cols, u, start, deg, a.cta_th, a.wave_th, s.g, sink);
        lq_flush(s.qv, out, LQ_CAP / 2);
    }
    lq_flush(s.qv, out, 1);
    if (COUNT_EDGES) wg_add_u64(eacc, &a.ctr->level_edges[(ph + 1) & 3], s);
    if (__syncthreads_or(found) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)depth_next);
}

// Contract-expand (edge frontier -> edge frontier).  Incoming (vertex,
// predecessor) pairs are culled (bitmask, history hash, exact claim) and
// labelled at contraction (depth = depth_here), then the survivors'
// neighbour lists are expanded into the outgoing edge frontier
// (contract_expand_atomic/cta.cuh:573-709).
__device__ void ce_phase(const BfsArgs& a, int ph, int32_t depth_here, LevelShared& s) {
    const uint32_t n_in = phase_input_len(a, ph);
    phase_begin(a, ph, n_in);
    const int32_t* qk = a.qk[ph & 1];
    const int32_t* qv = a.qv[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    bool found = false;
    for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_in; base += gridDim.x * BFS_THREADS) {
        const uint32_t i = base + threadIdx.x;
        int32_t v = -1, start = 0, deg = 0;
        if (i < n_in) {
            v = qk[i];
            const int32_t p = qv[i];
            bool keep = valid_vertex(a, v);
            if (keep) keep = !visited_test(a.visited, v);
            if (keep) keep = !history_seen(s.hist, v);
            if (keep) keep = claim_exact(a.visited, v);
            if (keep) {
                a.depth[v] = depth_here;
                a.pred[v] = p;
                found = true;
                start = a.offs[v];
                deg = a.offs[v + 1] - start;
            }
        }
        auto sink = [&](bool valid, int32_t nbr, int32_t src) { lq_append(s.qp, out, valid, nbr, src); };
        tiered_gather(a.cols, v, start, deg, a.cta_th, a.wave_th, s.g, sink);
        lq_flush(s.qp, out, LQ_CAP / 2);
    }
    lq_flush(s.qp, out, 1);
    if (__syncthreads_or(found) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)depth_here);
}

// Two-phase contract (edge frontier -> vertex frontier + labels),
// two_phase/contract_atomic/cta.cuh:494-572.
__device__ void tp_contract_phase(const BfsArgs& a, int ph, int32_t depth_here, LevelShared& s) {
    const uint32_t n_in = phase_input_len(a, ph);
    phase_begin(a, ph, n_in);
    const int32_t* qk = a.qk[ph & 1];
    const int32_t* qv = a.qv[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    bool found = false;
    for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_in; base += gridDim.x * BFS_THREADS) {
        const uint32_t i = base + threadIdx.x;
        bool keep = false;
        int32_t v = -1;
        if (i < n_in) {
            v = qk[i];
            const int32_t p = qv[i];
            keep = valid_vertex(a, v);
            if (keep) keep = !visited_test(a.visited, v);
            if (keep) keep = !history_seen(s.hist, v);
            if (keep) keep = claim_exact(a.visited, v);
            if (keep) {
                a.depth[v] = depth_here;
                a.pred[v] = p;
                found = true;
            }
        }
        lq_append(s.qv, out, keep, v, 0);
        lq_flush(s.qv, out, LQ_CAP / 2);
    }
    lq_flush(s.qv, out, 1);
    if (__syncthreads_or(found) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)depth_here);
}

// Two-phase expand (vertex frontier -> edge frontier of (neighbour, pred)),
// two_phase/expand_atomic/cta.cuh:606-719.  No culling.
__device__ void tp_expand_phase(const BfsArgs& a, int ph, LevelShared& s) {
    const uint32_t n_in = phase_input_len(a, ph);
    phase_begin(a, ph, n_in);
    const int32_t* qin = a.qk[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_in; base += gridDim.x * BFS_THREADS) {
        const uint32_t i = base + threadIdx.x;
        int32_t u = -1, start = 0, deg = 0;
        if (i < n_in) {
            u = qin[i];
            if (valid_vertex(a, u)) {
                start = a.offs[u];
                deg = a.offs[u + 1] - start;
            }
        }
        auto sink = [&](bool valid, int32_t nbr, int32_t src) { lq_append(s.qp, out, valid, nbr, src); };
        tiered_gather(a.cols, u, start, deg, a.cta_th, a.wave_th, s.g, sink);
        lq_flush(s.qp, out, LQ_CAP / 2);
    }
    lq_flush(s.qp, out, 1);
}

// Two-phase filter (edge frontier -> edge frontier), bitmask and history
// culls only, no labelling (two_phase/filter_atomic/cta.cuh:308-388).
__device__ void tp_filter_phase(const BfsArgs& a, int ph, LevelShared& s) {
    const uint32_t n_in = phase_input_len(a, ph);
    phase_begin(a, ph, n_in);
    const int32_t* qk = a.qk[ph & 1];
    const int32_t* qv = a.qv[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_in; base += gridDim.x * BFS_THREADS) {
        const uint32_t i = base + threadIdx.x;
        bool keep = false;
        int32_t v = -1, p = 0;
        if (i < n_in) {
            v = qk[i];
            p = qv[i];
            keep = valid_vertex(a, v);
            if (keep) keep = !visited_test(a.visited, v);
            if (keep) keep = !history_seen(s.hist, v);
        }
        lq_append(s.qp, out, keep, v, p);
        lq_flush(s.qp, out, LQ_CAP / 2);
    }
    lq_flush(s.qp, out, 1);
}

// Bottom-up step of the direction-optimizing traversal (no counterpart in
// the reference).  Every unvisited vertex scans its neighbours for one in the
// current frontier bitmap and stops at the first hit.  A wave owns the two
// bitmap words of its 64 vertices, so visited/next words are written without
// atomics.  Discovered vertices are also appended to the vertex queue so a
// top-down step can follow directly.
__device__ void bu_phase(const BfsArgs& a, int ph, int32_t depth_next, const uint32_t* cur, uint32_t* next,
                         LevelShared& s) {
    phase_begin(a, ph, 0);
    const GlobalQueue out = out_queue(a, ph);
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    const int64_t wave0 = (int64_t)blockIdx.x * BFS_WAVES + wid;
    const int64_t nwaves = (int64_t)gridDim.x * BFS_WAVES;
    bool found_any = false;
    unsigned long long eacc = 0;
    for (int64_t vb = wave0 * kWave; vb < a.n; vb += nwaves * kWave) {
        const int32_t v = (int32_t)(vb + lane);
        const uint32_t vw = (v < a.n) ? a.visited[v >> 5] : 0xFFFFFFFFu;
        const bool active = (v < a.n) && !((vw >> (v & 31)) & 1u);
        int32_t parent = -1;
        int32_t deg = 0;
        if (active) {
            const int32_t e0 = a.offs[v], e1 = a.offs[v + 1];
            deg = e1 - e0;
            for (int32_t e = e0; e < e1; ++e) {
                const int32_t u = a.cols[e];
                if ((cur[u >> 5] >> (u & 31)) & 1u) {
                    parent = u;
                    break;
                }
            }
        }
        const bool found = parent >= 0;
        if (found) {
            a.depth[v] = depth_next;
            a.pred[v] = parent;
            found_any = true;
            eacc += (unsigned long long)deg;
        }
        const uint64_t fm = __ballot(found);
        if ((lane & 31) == 0) {
            const int64_t w = (vb >> 5) + (lane >> 5);
            if (w < a.words) {
                const uint32_t bits = (lane == 0) ? (uint32_t)fm : (uint32_t)(fm >> 32);
                next[w] = bits;
                if (bits) a.visited[w] = vw | bits;
            }
        }
        lq_append(s.qv, out, found, v, 0);
    }
    lq_flush(s.qv, out, 1);
    wg_add_u64(eacc, &a.ctr->level_edges[(ph + 1) & 3], s);
    if (__syncthreads_or(found_any) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)depth_next);
}

// ===========================================================================
// Kernels.
// ===========================================================================

enum SeedMode { SEED_VERTEX = 0, SEED_EDGE = 1 };

// Resets the counters and enqueues the source (the iteration-0 prologue of
// contract_expand_atomic/kernel.cuh:307-332 and expand_contract_atomic/
// kernel.cuh:305-319).  Vertex mode labels the source at once (depth 0,
// pred -2); edge mode enqueues (src, -2) for labelling at contraction.
__global__ void k_seed(BfsArgs a, int ph, int32_t src, int mode) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DevCounters* c = a.ctr;
    for (int i = 0; i < 4; ++i) {
        c->qlen[i] = 0;
        c->level_edges[i] = 0;
        c->level_verts[i] = 0;
    }
    c->overflow = 0;
    c->last_level = 0;
    c->levels_run = 0;
    c->saturated = 0;
    c->queued = 0;
    c->next_edges = 0;
    c->next_edges_prev = 0;
    c->bar_count = 0;
    c->bar_gen = 0;
    c->bar_abort = 0;
    c->corrupt = 0;
    *a.max_depth = 0;
    a.qk[ph & 1][0] = src;
    if (mode == SEED_VERTEX) {
        a.visited[src >> 5] |= 1u << (src & 31);
        a.depth[src] = 0;
        a.pred[src] = -2;
        c->level_edges[ph & 3] = (unsigned long long)(a.offs[src + 1] - a.offs[src]);
    } else {
        a.qv[ph & 1][0] = -2;
    }
    c->qlen[ph & 3] = 1;
}

__device__ __forceinline__ void stats_begin(unsigned long long& t0, const BfsArgs& a) {
    if (a.wg_stats && threadIdx.x == 0) t0 = wg_clock();
}
__device__ __forceinline__ void stats_end(unsigned long long t0, const BfsArgs& a, int slot) {
    if (a.wg_stats && threadIdx.x == 0) {
        const unsigned long long t1 = wg_clock();
        atomic_add_u64(&a.wg_stats[0], t1 - t0);
        // per-launch span: [2 + 2*slot] = min start, [3 + 2*slot] = max end
        __hip_atomic_fetch_min(&a.wg_stats[2 + 2 * slot], t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_max(&a.wg_stats[3 + 2 * slot], t1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

enum PhaseKind { PK_TD = 0, PK_TD_COUNT = 1, PK_CE = 2, PK_TP_CONTRACT = 3, PK_TP_EXPAND = 4, PK_TP_FILTER = 5 };

template <int KIND>
__global__ __launch_bounds__(BFS_THREADS) void k_phase(BfsArgs a, int ph, int32_t depth, int slot) {
    __shared__ LevelShared s;
    unsigned long long t0 = 0;
    stats_begin(t0, a);
    shared_init(s);
    if (KIND == PK_TD) td_phase<false>(a, ph, depth, s);
    if (KIND == PK_TD_COUNT) td_phase<true>(a, ph, depth, s);
    if (KIND == PK_CE) ce_phase(a, ph, depth, s);
    if (KIND == PK_TP_CONTRACT) tp_contract_phase(a, ph, depth, s);
    if (KIND == PK_TP_EXPAND) tp_expand_phase(a, ph, s);
    if (KIND == PK_TP_FILTER) tp_filter_phase(a, ph, s);
    stats_end(t0, a, slot);
}

__global__ __launch_bounds__(BFS_THREADS) void k_bottom_up(BfsArgs a, int ph, int32_t depth_next, int cur_sel,
                                                           int slot) {
    __shared__ LevelShared s;
    unsigned long long t0 = 0;
    stats_begin(t0, a);
    shared_init(s);
    bu_phase(a, ph, depth_next, a.fbm[cur_sel], a.fbm[cur_sel ^ 1], s);
    stats_end(t0, a, slot);
}

// Frontier queue (slot ph) -> frontier bitmap (top-down to bottom-up switch).
__global__ void k_queue_to_bitmap(BfsArgs a, int ph, uint32_t* bm) {
    const uint32_t n_in = ld_relaxed(&a.ctr->qlen[ph & 3]);
    const int32_t* q = a.qk[ph & 1];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_in; i += gridDim.x * blockDim.x) {
        const int32_t v = q[i];
        if (valid_vertex(a, v)) atomic_or_u32(&bm[v >> 5], 1u << (v & 31));
    }
}

// Persistent kernels: all levels in one cooperative launch with a software
// grid barrier between phases (contract_expand_atomic::KernelGlobalBarrier,
// kernel.cuh:281-554; two_phase::Kernel, two_phase/kernel.cuh:60-313).
// Every workgroup reads the same queue length after each barrier, so all
// leave the loop together.  With `sat_limit` set, the kernel quits when a
// frontier exceeds it (SATURATION_QUIT) and records where it stopped.
enum FusedKind { FK_TD = 0, FK_CE = 1, FK_TP = 2 };

template <int KIND>
__global__ __launch_bounds__(BFS_THREADS) void k_fused(BfsArgs a, int ph, int32_t depth, int max_levels,
                                                       uint32_t sat_limit) {
    __shared__ LevelShared s;
    unsigned long long t0 = 0, busy = 0;
    if (a.wg_stats && threadIdx.x == 0) t0 = wg_clock();
    shared_init(s);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctr->saturated = 0;
    int level = 0;
    for (; level < max_levels; ++level) {
        if (ld_relaxed(&a.ctr->bar_abort)) break;
        const uint32_t n_in = ld_relaxed(&a.ctr->qlen[ph & 3]);
        if (n_in == 0 || n_in > a.qcap[ph & 1]) break;
        if (sat_limit && n_in > sat_limit) {
            if (blockIdx.x == 0 && threadIdx.x == 0) a.ctr->saturated = 1;
            break;
        }
        unsigned long long tb = 0;
        if (a.wg_stats && threadIdx.x == 0) tb = wg_clock();
        if (KIND == FK_TD) {
            td_phase<false>(a, ph, depth + 1, s);
            ph += 1;
        } else if (KIND == FK_CE) {
            ce_phase(a, ph, depth, s);
            ph += 1;
        } else {
            tp_contract_phase(a, ph, depth, s);
            if (a.wg_stats && threadIdx.x == 0) busy += wg_clock() - tb;
            grid_barrier(&a.ctr->bar_count, &a.ctr->bar_gen, gridDim.x, &a.ctr->bar_abort);
            if (ld_relaxed(&a.ctr->bar_abort)) break;
            if (a.wg_stats && threadIdx.x == 0) tb = wg_clock();
            tp_expand_phase(a, ph + 1, s);
            ph += 2;
        }
        depth += 1;
        if (a.wg_stats && threadIdx.x == 0) busy += wg_clock() - tb;
        grid_barrier(&a.ctr->bar_count, &a.ctr->bar_gen, gridDim.x, &a.ctr->bar_abort);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctr->levels_run = (uint32_t)level;
        a.ctr->last_level = (uint32_t)ph;
    }
    if (a.wg_stats && threadIdx.x == 0) {
        atomic_add_u64(&a.wg_stats[0], busy);
        atomic_add_u64(&a.wg_stats[1], wg_clock() - t0);
    }
}

// Global-barrier benchmark kernel (test/gbarrier/test_gbarrier.cu:43-50).
__global__ __launch_bounds__(BFS_THREADS) void k_gbarrier_bench(uint32_t* bar, int iterations) {
    for (int i = 0; i < iterations; ++i) {
        if (ld_relaxed(&bar[2])) break;
        grid_barrier(&bar[0], &bar[1], gridDim.x, &bar[2]);
    }
}

// Cross-workgroup visibility check of the grid barrier: each workgroup writes
// a slab of words, passes the barrier, then compares three other
// workgroups' slabs with the expected pattern.  Every index is fixed by
// (block, lane) and in bounds; loaded values are only compared.
__device__ __forceinline__ uint32_t gb_pattern(uint32_t it, uint32_t b, uint32_t j) {
    uint32_t x = it * 0x9E3779B1u ^ (b * 0x85EBCA6Bu) ^ (j * 0xC2B2AE35u);
    x ^= x >> 15;
    x *= 0x2C1B3C6Du;
    x ^= x >> 12;
    return x | 1u;
}

__global__ __launch_bounds__(BFS_THREADS) void k_gbarrier_check(uint32_t* bar, uint32_t* data, int words,
                                                                int iterations, unsigned long long* mismatches) {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    unsigned long long bad = 0;
    for (int it = 0; it < iterations; ++it) {
        if (ld_relaxed(&bar[2])) break;
        for (int j = threadIdx.x; j < words; j += BFS_THREADS) data[(size_t)b * words + j] = gb_pattern(it, b, j);
        grid_barrier(&bar[0], &bar[1], G, &bar[2]);
        const uint32_t others[3] = {(b + 1) % G, (b + G / 2 + 1) % G, (b + G - 1) % G};
        for (int k = 0; k < 3; ++k) {
            const uint32_t o = others[k];
            for (int j = threadIdx.x; j < words; j += BFS_THREADS)
                bad += (data[(size_t)o * words + j] != gb_pattern(it, o, j)) ? 1u : 0u;
        }
        grid_barrier(&bar[0], &bar[1], G, &bar[2]);
    }
    if (bad) atomic_add_u64(mismatches, bad);
}

}  // namespace b40c

// ===========================================================================
// Host objects.
// ===========================================================================

struct b40c_problem {
    b40c_graph* g = nullptr;
    hipStream_t stream = nullptr;
    int32_t flags = 0;
    b40c::DeviceBuffer depth, pred, visited, fbm0, fbm1, counters, maxdepth;
    b40c::DeviceBuffer qk[2], qv[2];
    size_t qcap[2] = {0, 0};
    size_t vcap[2] = {0, 0};
    int32_t frontier_type = -1;
    bool ready = false;
    b40c::DevCounters* h_ctr = nullptr;  // pinned readback
};

struct b40c_enactor {
    int32_t strategy = B40C_HYBRID;
    bool instrument = false;
    b40c_tuning tuning{};
    int64_t total_queued = 0;
    int32_t search_depth = 0;
    double avg_duty = 0.0;
    b40c::DeviceBuffer wg_stats;
    int launch_slots = 0;
    int fused_grid_cache = 0;
};

namespace b40c {
namespace {

constexpr int kMaxLaunchSlots = 4096;

// The persistent (cooperative, grid-barrier) kernels are switched off: a GPU
// memory fault occurred in a run that exercised them and its cause has not
// been shown yet.  Fused modes return B40C_ERR_UNSUPPORTED; the default modes
// use one launch per level instead.
constexpr bool kFusedEnabled = false;

b40c_tuning default_tuning(const DeviceInfo& info) {
    b40c_tuning t{};
    t.grid_size = info.num_cus * 4;
    t.cta_gather_threshold = 4096;
    t.wave_gather_threshold = 64;
    t.levels_per_sync = 4;
    t.saturation_quit = 512;
    t.do_alpha = 14;
    t.do_beta = 24;
    t.persistent_wgs_per_cu = 2;
    return t;
}

void merge_tuning(b40c_tuning& dst, const b40c_tuning& src) {
    if (src.grid_size > 0) dst.grid_size = src.grid_size;
    if (src.cta_gather_threshold > 0) dst.cta_gather_threshold = src.cta_gather_threshold;
    if (src.wave_gather_threshold > 0) dst.wave_gather_threshold = src.wave_gather_threshold;
    if (src.levels_per_sync > 0) dst.levels_per_sync = src.levels_per_sync;
    if (src.saturation_quit > 0) dst.saturation_quit = src.saturation_quit;
    if (src.do_alpha > 0) dst.do_alpha = src.do_alpha;
    if (src.do_beta > 0) dst.do_beta = src.do_beta;
    if (src.persistent_wgs_per_cu > 0) dst.persistent_wgs_per_cu = src.persistent_wgs_per_cu;
}

int32_t frontier_type_of(int32_t strategy) {
    switch (strategy) {
        case B40C_EXPAND_CONTRACT:
        case B40C_PERSISTENT:
        case B40C_DIRECTION_OPTIMIZING:
            return B40C_VERTEX_FRONTIERS;
        case B40C_TWO_PHASE:
            return B40C_MIXED_FRONTIERS;
        default:
            return B40C_EDGE_FRONTIERS;
    }
}

BfsArgs make_args(b40c_problem* p, const b40c_enactor* e) {
    BfsArgs a{};
    a.offs = p->g->d_offsets();
    a.cols = p->g->d_cols();
    a.depth = p->depth.as<int32_t>();
    a.pred = p->pred.as<int32_t>();
    a.visited = p->visited.as<uint32_t>();
    a.fbm[0] = p->fbm0.as<uint32_t>();
    a.fbm[1] = p->fbm1.as<uint32_t>();
    for (int i = 0; i < 2; ++i) {
        a.qk[i] = p->qk[i].as<int32_t>();
        a.qv[i] = p->qv[i].as<int32_t>();
        a.qcap[i] = (uint32_t)std::min<size_t>(p->qcap[i], 0xFFFFFFFFull);
    }
    a.ctr = p->counters.as<DevCounters>();
    a.max_depth = p->maxdepth.as<uint32_t>();
    a.wg_stats = e->instrument ? e->wg_stats.as<unsigned long long>() : nullptr;
    a.n = p->g->nodes;
    a.words = (p->g->nodes + 31) / 32;
    // Clamped so the fine tier's workgroup scan (< 256 x threshold) fits int32.
    a.cta_th = std::min(std::max(1, e->tuning.cta_gather_threshold), 1 << 20);
    a.wave_th = std::min(std::max(1, e->tuning.wave_gather_threshold), 1 << 20);
    return a;
}

b40c_status read_counters(b40c_problem* p) {
    B40C_HIP(hipMemcpyAsync(p->h_ctr, p->counters.ptr, sizeof(DevCounters), hipMemcpyDeviceToHost, p->stream));
    B40C_HIP(hipStreamSynchronize(p->stream));
    return B40C_OK;
}

struct Launcher {
    b40c_problem* p;
    b40c_enactor* e;
    BfsArgs a;
    unsigned grid;

    int next_slot() {
        if (!e->instrument) return 0;
        int s = e->launch_slots;
        if (s < kMaxLaunchSlots - 1) e->launch_slots++;
        return s;
    }
    b40c_status phase(int kind, int ph, int32_t depth) {
        const int slot = next_slot();
        switch (kind) {
            case PK_TD:
                hipLaunchKernelGGL(k_phase<PK_TD>, dim3(grid), dim3(BFS_THREADS), 0, p->stream, a, ph, depth, slot);
                break;
            case PK_TD_COUNT:
                hipLaunchKernelGGL(k_phase<PK_TD_COUNT>, dim3(grid), dim3(BFS_THREADS), 0, p->stream, a, ph, depth,
                                   slot);
                break;
            case PK_CE:
                hipLaunchKernelGGL(k_phase<PK_CE>, dim3(grid), dim3(BFS_THREADS), 0, p->stream, a, ph, depth, slot);
                break;
            case PK_TP_CONTRACT:
                hipLaunchKernelGGL(k_phase<PK_TP_CONTRACT>, dim3(grid), dim3(BFS_THREADS), 0, p->stream, a, ph, depth,
                                   slot);
                break;
            case PK_TP_EXPAND:
                hipLaunchKernelGGL(k_phase<PK_TP_EXPAND>, dim3(grid), dim3(BFS_THREADS), 0, p->stream, a, ph, depth,
                                   slot);
                break;
            case PK_TP_FILTER:
                hipLaunchKernelGGL(k_phase<PK_TP_FILTER>, dim3(grid), dim3(BFS_THREADS), 0, p->stream, a, ph, depth,
                                   slot);
                break;
        }
        B40C_HIP(hipGetLastError());
        return B40C_OK;
    }
    b40c_status seed(int ph, int32_t src, int mode) {
        hipLaunchKernelGGL(k_seed, dim3(1), dim3(64), 0, p->stream, a, ph, src, mode);
        B40C_HIP(hipGetLastError());
        return B40C_OK;
    }
    b40c_status fused(int kind, int ph, int32_t depth, int max_levels, uint32_t sat, unsigned fgrid) {
        if (!kFusedEnabled)
            return fail(B40C_ERR_UNSUPPORTED, "persistent (fused) kernels are disabled in this build");
        void* args[] = {&a, &ph, &depth, &max_levels, &sat};
        const void* fn = nullptr;
        if (kind == FK_TD) fn = (const void*)k_fused<FK_TD>;
        if (kind == FK_CE) fn = (const void*)k_fused<FK_CE>;
        if (kind == FK_TP) fn = (const void*)k_fused<FK_TP>;
        B40C_HIP(hipLaunchCooperativeKernel(fn, dim3(fgrid), dim3(BFS_THREADS), args, 0, p->stream));
        return B40C_OK;
    }
};

// Largest co-resident grid for a fused kernel, capped by the tuning.
b40c_status fused_grid(b40c_enactor* e, int kind, int max_grid, unsigned& out) {
    DeviceInfo info;
    B40C_TRY(device_info(info));
    int per_cu = 0;
    const void* fn = (kind == FK_TD) ? (const void*)k_fused<FK_TD>
                     : (kind == FK_CE) ? (const void*)k_fused<FK_CE>
                                       : (const void*)k_fused<FK_TP>;
    B40C_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BFS_THREADS, 0));
    if (per_cu < 1) return fail(B40C_ERR_UNSUPPORTED, "fused kernel cannot be resident");
    per_cu = std::min(per_cu, std::max(1, e->tuning.persistent_wgs_per_cu));
    int g = per_cu * info.num_cus;
    if (max_grid > 0) g = std::min(g, max_grid);
    out = (unsigned)std::max(1, g);
    return B40C_OK;
}

// ---- strategies --------------------------------------------------------------

// Host level loop: launches `k` phases at a time, then checks the queue
// length once (device-side termination makes surplus launches no-ops).
template <typename StepFn>
b40c_status level_loop(b40c_problem* p, b40c_enactor* e, int& ph, StepFn step, int phases_per_level) {
    int k = std::max(1, e->tuning.levels_per_sync);
    while (true) {
        for (int j = 0; j < k; ++j) B40C_TRY(step());
        B40C_TRY(read_counters(p));
        if (p->h_ctr->overflow) break;
        if (p->h_ctr->qlen[ph & 3] == 0) break;
        k = std::min(k * 2, 64);
    }
    (void)phases_per_level;
    return B40C_OK;
}

b40c_status enact_expand_contract(Launcher& L, int32_t src, bool fused) {
    int ph = 0;
    B40C_TRY(L.seed(ph, src, SEED_VERTEX));
    if (fused) {
        unsigned fg = 0;
        B40C_TRY(fused_grid(L.e, FK_TD, 0, fg));
        if (L.grid < fg) fg = L.grid;
        return L.fused(FK_TD, ph, 0, 1 << 30, 0, fg);
    }
    int32_t depth = 0;
    return level_loop(
        L.p, L.e, ph,
        [&]() -> b40c_status {
            B40C_TRY(L.phase(PK_TD, ph, depth + 1));
            ph += 1;
            depth += 1;
            return B40C_OK;
        },
        1);
}

b40c_status enact_contract_expand(Launcher& L, int32_t src, bool fused) {
    int ph = 0;
    B40C_TRY(L.seed(ph, src, SEED_EDGE));
    if (fused) {
        unsigned fg = 0;
        B40C_TRY(fused_grid(L.e, FK_CE, 0, fg));
        if (L.grid < fg) fg = L.grid;
        return L.fused(FK_CE, ph, 0, 1 << 30, 0, fg);
    }
    int32_t depth = 0;
    return level_loop(
        L.p, L.e, ph,
        [&]() -> b40c_status {
            B40C_TRY(L.phase(PK_CE, ph, depth));
            ph += 1;
            depth += 1;
            return B40C_OK;
        },
        1);
}

b40c_status enact_two_phase(Launcher& L, int32_t src, bool fused) {
    // Mixed frontiers: buffer 1 holds edge frontiers, buffer 0 vertex
    // frontiers, so contraction runs on odd phases.
    int ph = 1;
    B40C_TRY(L.seed(ph, src, SEED_EDGE));
    if (fused) {
        unsigned fg = 0;
        B40C_TRY(fused_grid(L.e, FK_TP, 0, fg));
        if (L.grid < fg) fg = L.grid;
        return L.fused(FK_TP, ph, 0, 1 << 30, 0, fg);
    }
    int32_t depth = 0;
    return level_loop(
        L.p, L.e, ph,
        [&]() -> b40c_status {
            B40C_TRY(L.phase(PK_TP_CONTRACT, ph, depth));
            B40C_TRY(L.phase(PK_TP_EXPAND, ph + 1, depth));
            ph += 2;
            depth += 1;
            return B40C_OK;
        },
        2);
}

// Hybrid (enactor_hybrid.cuh:273-551): fused contract-expand while the edge
// frontier stays below grid * saturation_quit; when it saturates, the host
// runs filter -> contract -> expand steps until the filtered frontier is
// small again, then resumes the fused kernel.
b40c_status enact_hybrid(Launcher& L, int32_t src) {
    b40c_problem* p = L.p;
    b40c_enactor* e = L.e;
    int ph = 0;
    int32_t depth = 0;
    B40C_TRY(L.seed(ph, src, SEED_EDGE));
    unsigned fg = L.grid;
    if (kFusedEnabled) {
        B40C_TRY(fused_grid(e, FK_CE, 0, fg));
        if (L.grid < fg) fg = L.grid;
    }
    const uint32_t sat = (uint32_t)std::max<int64_t>(1, (int64_t)fg * std::max(1, e->tuning.saturation_quit));
    while (true) {
        if (kFusedEnabled) {
            B40C_TRY(L.fused(FK_CE, ph, depth, 1 << 30, sat, fg));
            B40C_TRY(read_counters(p));
            const uint32_t levels = p->h_ctr->levels_run;
            ph = (int)p->h_ctr->last_level;
            depth += (int32_t)levels;
            if (p->h_ctr->overflow || !p->h_ctr->saturated) break;
        } else {
            // Narrow levels: one contract-expand launch per level.
            bool saturated = false;
            while (true) {
                B40C_TRY(L.phase(PK_CE, ph, depth));
                ph += 1;
                depth += 1;
                B40C_TRY(read_counters(p));
                const uint32_t n = p->h_ctr->qlen[ph & 3];
                if (p->h_ctr->overflow || n == 0) break;
                if (n > sat) {
                    saturated = true;
                    break;
                }
            }
            if (!saturated) break;
        }
        // Wide levels: filter -> contract -> expand, one launch each.
        while (true) {
            B40C_TRY(L.phase(PK_TP_FILTER, ph, depth));
            B40C_TRY(L.phase(PK_TP_CONTRACT, ph + 1, depth));
            B40C_TRY(L.phase(PK_TP_EXPAND, ph + 2, depth));
            ph += 3;
            depth += 1;
            B40C_TRY(read_counters(p));
            if (p->h_ctr->overflow) break;
            const uint32_t n = p->h_ctr->qlen[ph & 3];
            if (n == 0 || n <= sat) break;
        }
        if (p->h_ctr->overflow || p->h_ctr->qlen[ph & 3] == 0) break;
    }
    return B40C_OK;
}

// Direction-optimizing traversal (Beamer, Asanovic, Patterson, SC 2012): top
// down while the frontier's edges are few, bottom up while the frontier is a
// large share of the graph.  m_f = degree sum of the frontier, m_u = edges of
// unvisited vertices; switch down->up when m_f > m_u / alpha, up->down when
// n_f < n / beta.
b40c_status enact_direction_optimizing(Launcher& L, int32_t src) {
    b40c_problem* p = L.p;
    b40c_enactor* e = L.e;
    const int64_t n = p->g->nodes;
    const int64_t E = p->g->edges;
    const int alpha = std::max(1, e->tuning.do_alpha);
    const int beta = std::max(1, e->tuning.do_beta);
    int ph = 0;
    int32_t depth = 0;
    B40C_TRY(L.seed(ph, src, SEED_VERTEX));
    int64_t m_visited = 0;  // degree sum of vertices visited so far (excluding frontier)
    int64_t m_f = 0, n_f = 1;
    B40C_TRY(read_counters(p));
    m_f = (int64_t)p->h_ctr->level_edges[ph & 3];
    bool bottom_up = false;
    int cur_sel = 0;
    while (n_f > 0) {
        const int64_t m_u = E - m_visited - m_f;
        if (!bottom_up && m_f > m_u / alpha) {
            bottom_up = true;
            cur_sel = 0;
            B40C_HIP(hipMemsetAsync(p->fbm0.ptr, 0, p->fbm0.bytes, p->stream));
            hipLaunchKernelGGL(k_queue_to_bitmap, dim3(L.grid), dim3(256), 0, p->stream, L.a, ph,
                               p->fbm0.as<uint32_t>());
            B40C_HIP(hipGetLastError());
        } else if (bottom_up && n_f < n / beta) {
            bottom_up = false;
        }
        m_visited += m_f;
        if (bottom_up) {
            const int slot = L.next_slot();
            hipLaunchKernelGGL(k_bottom_up, dim3(L.grid), dim3(BFS_THREADS), 0, p->stream, L.a, ph, depth + 1,
                               cur_sel, slot);
            B40C_HIP(hipGetLastError());
            cur_sel ^= 1;
        } else {
            B40C_TRY(L.phase(PK_TD_COUNT, ph, depth + 1));
        }
        ph += 1;
        depth += 1;
        B40C_TRY(read_counters(p));
        if (p->h_ctr->overflow) break;
        n_f = p->h_ctr->qlen[ph & 3];
        m_f = (int64_t)p->h_ctr->level_edges[ph & 3];
    }
    return B40C_OK;
}

}  // namespace
}  // namespace b40c

using namespace b40c;

extern "C" {

b40c_status b40c_problem_create_from_graph(b40c_graph* g, int32_t flags, b40c_problem** out) {
    if (!g || !out) return fail(B40C_ERR_INVALID_ARG, "b40c_problem_create_from_graph: null argument");
    if (flags != 0) return fail(B40C_ERR_INVALID_ARG, "b40c_problem_create_from_graph: unknown flags %d", flags);
    b40c_problem* p = new b40c_problem();
    p->flags = flags;
    hipError_t err = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete p;
        return fail(B40C_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(err));
    }
    err = hipHostMalloc((void**)&p->h_ctr, sizeof(DevCounters), hipHostMallocDefault);
    if (err != hipSuccess) {
        (void)hipStreamDestroy(p->stream);
        delete p;
        return fail(B40C_ERR_HIP, "hipHostMalloc failed: %s", hipGetErrorString(err));
    }
    graph_retain(g);
    p->g = g;
    const size_t n = (size_t)g->nodes;
    const size_t words = (n + 31) / 32;
    b40c_status s = p->depth.alloc(n * 4);
    if (s == B40C_OK) s = p->pred.alloc(n * 4);
    if (s == B40C_OK) s = p->visited.alloc(words * 4);
    if (s == B40C_OK) s = p->fbm0.alloc(words * 4);
    if (s == B40C_OK) s = p->fbm1.alloc(words * 4);
    if (s == B40C_OK) s = p->counters.alloc(sizeof(DevCounters));
    if (s == B40C_OK) s = p->maxdepth.alloc(256);
    if (s != B40C_OK) {
        b40c_problem_destroy(p);
        return s;
    }
    *out = p;
    return B40C_OK;
}

b40c_status b40c_problem_create(int32_t nodes, int64_t edges, const int32_t* column_indices,
                                const int32_t* row_offsets, int32_t flags, b40c_problem** out) {
    b40c_graph* g = nullptr;
    B40C_TRY(b40c_graph_from_host_csr(nodes, edges, row_offsets, column_indices, &g));
    b40c_status s = b40c_problem_create_from_graph(g, flags, out);
    graph_release(g);  // the problem holds its own reference
    return s;
}

b40c_status b40c_problem_reset(b40c_problem* p, int32_t frontier_type, double queue_sizing) {
    if (!p) return fail(B40C_ERR_INVALID_ARG, "b40c_problem_reset: null problem");
    if (frontier_type < B40C_VERTEX_FRONTIERS || frontier_type > B40C_MIXED_FRONTIERS)
        return fail(B40C_ERR_INVALID_ARG, "b40c_problem_reset: bad frontier type %d", frontier_type);
    if (queue_sizing <= 0.0) queue_sizing = 1.3;
    const double n = (double)p->g->nodes, m = (double)std::max<int64_t>(p->g->edges, 1);
    // Queue sizes per frontier type (csr_problem.cuh:602-641).
    size_t want[2], wantv[2];
    if (frontier_type == B40C_VERTEX_FRONTIERS) {
        want[0] = want[1] = (size_t)(n * queue_sizing);
        wantv[0] = wantv[1] = 0;
    } else if (frontier_type == B40C_EDGE_FRONTIERS) {
        want[0] = want[1] = (size_t)(m * queue_sizing);
        wantv[0] = wantv[1] = want[0];
    } else {
        want[0] = (size_t)(n * queue_sizing);
        want[1] = (size_t)(m * queue_sizing);
        wantv[0] = 0;
        wantv[1] = want[1];
    }
    for (int i = 0; i < 2; ++i) {
        want[i] = std::max<size_t>(want[i], 1);
        want[i] = std::min<size_t>(want[i], 0xFFFFFFF0ull);
        wantv[i] = std::min<size_t>(wantv[i], 0xFFFFFFF0ull);
        if (p->qk[i].bytes < want[i] * 4) B40C_TRY(p->qk[i].alloc(want[i] * 4));
        if (wantv[i] && p->qv[i].bytes < wantv[i] * 4) B40C_TRY(p->qv[i].alloc(wantv[i] * 4));
        // Capacity is what this reset asked for, so a small queue_sizing
        // really limits the frontier (and can overflow).
        p->qcap[i] = want[i];
        p->vcap[i] = wantv[i];
    }
    p->frontier_type = frontier_type;
    const size_t nb = (size_t)p->g->nodes * 4;
    B40C_HIP(hipMemsetAsync(p->depth.ptr, 0xFF, nb, p->stream));
    B40C_HIP(hipMemsetAsync(p->pred.ptr, 0xFF, nb, p->stream));
    B40C_HIP(hipMemsetAsync(p->visited.ptr, 0, p->visited.bytes, p->stream));
    p->ready = true;
    return B40C_OK;
}

b40c_status b40c_problem_extract(b40c_problem* p, int32_t* depth, int32_t* pred) {
    if (!p) return fail(B40C_ERR_INVALID_ARG, "b40c_problem_extract: null problem");
    const size_t nb = (size_t)p->g->nodes * 4;
    if (depth) B40C_HIP(hipMemcpyAsync(depth, p->depth.ptr, nb, hipMemcpyDeviceToHost, p->stream));
    if (pred) B40C_HIP(hipMemcpyAsync(pred, p->pred.ptr, nb, hipMemcpyDeviceToHost, p->stream));
    B40C_HIP(hipStreamSynchronize(p->stream));
    return B40C_OK;
}

void* b40c_problem_stream(b40c_problem* p) { return p ? (void*)p->stream : nullptr; }

b40c_status b40c_problem_device_labels(b40c_problem* p, int32_t** d_depth, int32_t** d_pred) {
    if (!p) return fail(B40C_ERR_INVALID_ARG, "b40c_problem_device_labels: null problem");
    if (d_depth) *d_depth = p->depth.as<int32_t>();
    if (d_pred) *d_pred = p->pred.as<int32_t>();
    return B40C_OK;
}

void b40c_problem_destroy(b40c_problem* p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    p->depth.release();
    p->pred.release();
    p->visited.release();
    p->fbm0.release();
    p->fbm1.release();
    p->counters.release();
    p->maxdepth.release();
    for (int i = 0; i < 2; ++i) {
        p->qk[i].release();
        p->qv[i].release();
    }
    if (p->h_ctr) (void)hipHostFree(p->h_ctr);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    graph_release(p->g);
    delete p;
}

b40c_status b40c_enactor_create(int32_t strategy, int32_t instrument, b40c_enactor** out) {
    if (!out) return fail(B40C_ERR_INVALID_ARG, "b40c_enactor_create: null out");
    if (strategy < B40C_EXPAND_CONTRACT || strategy > B40C_DIRECTION_OPTIMIZING)
        return fail(B40C_ERR_INVALID_ARG, "b40c_enactor_create: unknown strategy %d", strategy);
    DeviceInfo info;
    B40C_TRY(device_info(info));
    if (info.arch.rfind("gfx950", 0) != 0)
        return fail(B40C_ERR_UNSUPPORTED, "b40c_enactor_create: device %s is not gfx950", info.arch.c_str());
    b40c_enactor* e = new b40c_enactor();
    e->strategy = strategy;
    e->instrument = instrument != 0;
    e->tuning = default_tuning(info);
    if (e->instrument) {
        b40c_status s = e->wg_stats.alloc((2 + 2 * kMaxLaunchSlots) * sizeof(unsigned long long));
        if (s != B40C_OK) {
            delete e;
            return s;
        }
    }
    *out = e;
    return B40C_OK;
}

b40c_status b40c_enactor_set_tuning(b40c_enactor* e, const b40c_tuning* t) {
    if (!e || !t) return fail(B40C_ERR_INVALID_ARG, "b40c_enactor_set_tuning: null argument");
    merge_tuning(e->tuning, *t);
    return B40C_OK;
}

b40c_status b40c_enactor_get_tuning(const b40c_enactor* e, b40c_tuning* t) {
    if (!e || !t) return fail(B40C_ERR_INVALID_ARG, "b40c_enactor_get_tuning: null argument");
    *t = e->tuning;
    return B40C_OK;
}

int32_t b40c_enactor_frontier_type(const b40c_enactor* e) { return e ? frontier_type_of(e->strategy) : -1; }

b40c_status b40c_enact_search(b40c_enactor* e, b40c_problem* p, int32_t src, int32_t max_grid_size, int32_t mode) {
    if (!e || !p) return fail(B40C_ERR_INVALID_ARG, "b40c_enact_search: null argument");
    if (!p->ready) return fail(B40C_ERR_INVALID_ARG, "b40c_enact_search: problem not reset since last search");
    if (src < 0 || src >= p->g->nodes)
        return fail(B40C_ERR_INVALID_ARG, "b40c_enact_search: source %d out of range [0,%d)", src, p->g->nodes);
    const int32_t need = frontier_type_of(e->strategy);
    const bool edge_ok = (p->frontier_type == B40C_EDGE_FRONTIERS);
    if (p->frontier_type != need && !(need == B40C_VERTEX_FRONTIERS && edge_ok) &&
        !(need == B40C_MIXED_FRONTIERS && edge_ok))
        return fail(B40C_ERR_INVALID_ARG, "b40c_enact_search: problem reset with frontier type %d, strategy needs %d",
                    p->frontier_type, need);
    if (mode < B40C_MODE_DEFAULT || mode > B40C_MODE_ITERATIVE)
        return fail(B40C_ERR_INVALID_ARG, "b40c_enact_search: bad mode %d", mode);
    p->ready = false;

    Launcher L{p, e, make_args(p, e), 0};
    int g = e->tuning.grid_size;
    if (max_grid_size > 0) g = std::min(g, max_grid_size);
    L.grid = (unsigned)std::max(1, g);
    e->launch_slots = 0;
    if (e->instrument) {
        std::vector<unsigned long long> init(2 + 2 * kMaxLaunchSlots);
        init[0] = init[1] = 0;
        for (int i = 0; i < kMaxLaunchSlots; ++i) {
            init[2 + 2 * i] = ~0ull;
            init[3 + 2 * i] = 0;
        }
        B40C_HIP(hipMemcpyAsync(e->wg_stats.ptr, init.data(), init.size() * 8, hipMemcpyHostToDevice, p->stream));
        B40C_HIP(hipStreamSynchronize(p->stream));
    }

    b40c_status s = B40C_OK;
    switch (e->strategy) {
        case B40C_EXPAND_CONTRACT:
            // Fused is the reference's default form for this strategy (test_bfs.cu:690-695).
            s = enact_expand_contract(L, src, mode == B40C_MODE_FUSED || (mode == B40C_MODE_DEFAULT && kFusedEnabled));
            break;
        case B40C_PERSISTENT:
            s = enact_expand_contract(L, src, mode != B40C_MODE_ITERATIVE);
            break;
        case B40C_CONTRACT_EXPAND:
            s = enact_contract_expand(L, src, mode == B40C_MODE_FUSED || (mode == B40C_MODE_DEFAULT && kFusedEnabled));
            break;
        case B40C_TWO_PHASE:
            s = enact_two_phase(L, src, mode == B40C_MODE_FUSED);
            break;
        case B40C_HYBRID:
            s = (mode == B40C_MODE_DEFAULT) ? enact_hybrid(L, src)
                                            : enact_contract_expand(L, src, mode == B40C_MODE_FUSED);
            break;
        case B40C_DIRECTION_OPTIMIZING:
            s = enact_direction_optimizing(L, src);
            break;
    }
    if (s != B40C_OK) return s;
    B40C_TRY(read_counters(p));
    uint32_t maxd = 0;
    B40C_HIP(hipMemcpy(&maxd, p->maxdepth.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost));
    e->total_queued = (int64_t)p->h_ctr->queued;
    e->search_depth = (int32_t)maxd + 1;
    e->avg_duty = 0.0;
    if (e->instrument) {
        std::vector<unsigned long long> st(2 + 2 * kMaxLaunchSlots);
        B40C_HIP(hipMemcpy(st.data(), e->wg_stats.ptr, st.size() * 8, hipMemcpyDeviceToHost));
        double life = (double)st[1];
        if (life == 0.0) {
            // Per-level kernels: lifetime = launch span x workgroups.
            for (int i = 0; i < e->launch_slots; ++i)
                if (st[3 + 2 * i] > st[2 + 2 * i]) life += (double)(st[3 + 2 * i] - st[2 + 2 * i]) * L.grid;
        }
        e->avg_duty = life > 0 ? (double)st[0] / life : 0.0;
    }
    if (p->h_ctr->bar_abort) return fail(B40C_ERR_HIP, "grid barrier wait timed out (grid not co-resident?)");
    if (p->h_ctr->corrupt) return fail(B40C_ERR_HIP, "frontier queue held a vertex id outside [0, n)");
    if (p->h_ctr->overflow)
        return fail(B40C_ERR_QUEUE_OVERFLOW, "Frontier queue overflow. Please increase queue-sizing factor.");
    return B40C_OK;
}

b40c_status b40c_get_statistics(b40c_enactor* e, int64_t* total_queued, int32_t* search_depth, double* avg_duty) {
    if (!e) return fail(B40C_ERR_INVALID_ARG, "b40c_get_statistics: null enactor");
    if (total_queued) *total_queued = e->total_queued;
    if (search_depth) *search_depth = e->search_depth;
    if (avg_duty) *avg_duty = e->avg_duty;
    return B40C_OK;
}

void b40c_enactor_destroy(b40c_enactor* e) { delete e; }

b40c_status b40c_gbarrier_benchmark(int32_t grid_size, int32_t iterations, double* ms_per_sync) {
    if (grid_size < 1 || iterations < 1 || !ms_per_sync)
        return fail(B40C_ERR_INVALID_ARG, "b40c_gbarrier_benchmark: bad arguments");
    DeviceInfo info;
    B40C_TRY(device_info(info));
    int per_cu = 0;
    B40C_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_gbarrier_bench, BFS_THREADS, 0));
    if (grid_size > per_cu * info.num_cus)
        return fail(B40C_ERR_INVALID_ARG, "b40c_gbarrier_benchmark: grid %d exceeds co-resident limit %d", grid_size,
                    per_cu * info.num_cus);
    DeviceBuffer bar;
    B40C_TRY(bar.alloc(256));
    B40C_HIP(hipMemset(bar.ptr, 0, 256));
    hipEvent_t e0, e1;
    B40C_HIP(hipEventCreate(&e0));
    B40C_HIP(hipEventCreate(&e1));
    uint32_t* barp = bar.as<uint32_t>();
    int iters = iterations;
    void* args[] = {&barp, &iters};
    B40C_HIP(hipEventRecord(e0, nullptr));
    B40C_HIP(hipLaunchCooperativeKernel((const void*)k_gbarrier_bench, dim3(grid_size), dim3(BFS_THREADS), args, 0,
                                        nullptr));
    B40C_HIP(hipEventRecord(e1, nullptr));
    B40C_HIP(hipEventSynchronize(e1));
    uint32_t aborted = 0;
    B40C_HIP(hipMemcpy(&aborted, barp + 2, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (aborted) return fail(B40C_ERR_HIP, "grid barrier wait timed out");
    float ms = 0;
    B40C_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_sync = ms / iterations;
    return B40C_OK;
}

b40c_status b40c_gbarrier_check(int32_t grid_size, int32_t iterations, int64_t* mismatches) {
    if (grid_size < 2 || iterations < 1 || !mismatches)
        return fail(B40C_ERR_INVALID_ARG, "b40c_gbarrier_check: bad arguments");
    DeviceInfo info;
    B40C_TRY(device_info(info));
    int per_cu = 0;
    B40C_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_gbarrier_check, BFS_THREADS, 0));
    if (grid_size > per_cu * info.num_cus)
        return fail(B40C_ERR_INVALID_ARG, "b40c_gbarrier_check: grid %d exceeds co-resident limit %d", grid_size,
                    per_cu * info.num_cus);
    const int words = 1024;
    DeviceBuffer bar, data, bad;
    B40C_TRY(bar.alloc(256));
    B40C_TRY(data.alloc((size_t)grid_size * words * sizeof(uint32_t)));
    B40C_TRY(bad.alloc(sizeof(unsigned long long)));
    B40C_HIP(hipMemset(bar.ptr, 0, 256));
    B40C_HIP(hipMemset(data.ptr, 0, data.bytes));
    B40C_HIP(hipMemset(bad.ptr, 0, sizeof(unsigned long long)));
    uint32_t* barp = bar.as<uint32_t>();
    uint32_t* datap = data.as<uint32_t>();
    unsigned long long* badp = bad.as<unsigned long long>();
    int w = words, iters = iterations;
    void* args[] = {&barp, &datap, &w, &iters, &badp};
    B40C_HIP(hipLaunchCooperativeKernel((const void*)k_gbarrier_check, dim3(grid_size), dim3(BFS_THREADS), args, 0,
                                        nullptr));
    B40C_HIP(hipDeviceSynchronize());
    uint32_t aborted = 0;
    unsigned long long h_bad = 0;
    B40C_HIP(hipMemcpy(&aborted, barp + 2, sizeof(uint32_t), hipMemcpyDeviceToHost));
    B40C_HIP(hipMemcpy(&h_bad, badp, sizeof(h_bad), hipMemcpyDeviceToHost));
    if (aborted) return fail(B40C_ERR_HIP, "grid barrier wait timed out");
    *mismatches = (int64_t)h_bad;
    return B40C_OK;
}

}  // extern "C"
