// Device side of the single-GPU BFS for gfx950: the level step bodies and
// every kernel's body as a __device__ function that takes its LDS block by
// reference.  bfs.hip wraps each body in a __global__ kernel that declares
// the LDS block; keeping the bodies here lets a CPU replay build (tests/
// cpu_replay) run the very same code with bounds-checked buffers.
//
// Reference counterparts (b40c):
//   level kernels             b40c/graph/bfs/{expand_contract_atomic,
//                             contract_expand_atomic, two_phase}/kernel.cuh
//   persistent kernels        contract_expand_atomic/kernel.cuh:281-554,
//                             two_phase/kernel.cuh:60-313
//
// Queue protocol: a "phase" is one kernel step that reads one frontier queue
// and writes the next.  Phase ph reads buffer ph&1 with its length in
// counter slot ph&3, appends to buffer (ph+1)&1 / slot (ph+1)&3 and clears
// slot (ph+2)&3.  Expand-contract and contract-expand use one phase per BFS
// level, two-phase uses two (contract, expand) or three (filter, contract,
// expand).  Depth labels are carried separately from the phase index.
//
// Grid tier: an expanding phase defers vertices whose degree reaches
// grid_th to a list (length slot ph&1); the k_heavy launch that follows the
// phase splits those adjacency lists into HEAVY_CHUNK-edge pieces spread over
// every workgroup, so one hub no longer serializes a level on one workgroup.
#pragma once

#include <climits>

#include "bfs_device.hpp"

namespace b40c {

// ===========================================================================
// Level bodies (device).
// ===========================================================================

// LDS of a level step.  Keys-only steps (expand-contract, two-phase
// contraction) and pair steps (contract-expand, two-phase filter) append to
// the same pair-capable queue, so a persistent kernel that alternates them
// keeps one valid count.
struct LevelShared {
    LdsPairQueue q;
    GatherSmem g;
    HistoryHash hist;
    unsigned long long red[BFS_WAVES];
    uint32_t ctl;  // persistent kernels: loop control, read by one thread per workgroup
};

__device__ __forceinline__ void shared_init(LevelShared& s) {
    if (threadIdx.x == 0) s.q.count = 0;
    gather_init(s.g);
    history_init(s.hist);
    __syncthreads();
}

// Frontier entries are validated before they are used as addresses: an id
// outside [0, n) is dropped and flagged, so a corrupt queue surfaces as an
// error status instead of an out-of-bounds access.
__device__ __forceinline__ bool valid_vertex(const BfsArgs& a, int32_t v) {
    if ((uint32_t)v < (uint32_t)a.n) return true;
    st_relaxed(&a.ctr->corrupt, 1u);
    return false;
}

__device__ __forceinline__ uint32_t phase_input_len(const BfsArgs& a, int ph) {
    uint32_t n = ld_relaxed(&a.ctr->qlen[ph & 3]);
    if (n > a.qcap[ph & 1]) n = 0;  // an overflowed queue counts as empty
    return n;
}

// A phase body's n_in argument: the input frontier's length when the caller
// already has it (the persistent kernels get it with the grid barrier), or
// this value to read it from the counters.
constexpr uint32_t kLenFromCounters = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t input_len(const BfsArgs& a, int ph, uint32_t n_in) {
    return n_in == kLenFromCounters ? phase_input_len(a, ph) : n_in;
}

__device__ __forceinline__ void phase_begin(const BfsArgs& a, int ph, uint32_t n_in) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st_relaxed(&a.ctr->qlen[(ph + 2) & 3], 0u);
        __hip_atomic_store(&a.ctr->heavy_alloc[(ph + 1) & 1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.ctr->level_edges[(ph + 2) & 3] = 0ull;
        if (n_in) atomic_add_u64(&a.ctr->queued, n_in);
    }
}

// Frontier entries per workgroup per step of a gathering phase.  Level
// launches take full tiles of BFS_THREADS.  A persistent launch spreads a
// frontier smaller than its grid's tiles evenly over all its workgroups: every level of a high-diameter graph is such a frontier, and
// spread out each workgroup gathers its few lists' edges in one round of
// the fine tier instead of a few workgroups gathering full tiles in several
// rounds, each a dependent chain (column, visited word, claiming atomic).
// Level launches run four times as many workgroups, and there the extra
// workgroups' same-address queue and depth atomics cost more than the
// rounds saved (200^3 lattice: 25 instead of 17 us per level).
__device__ __forceinline__ uint32_t entries_per_wg(const BfsArgs& a, uint32_t n_in, uint32_t nblocks) {
    if (!a.persistent || (uint64_t)n_in >= (uint64_t)nblocks * BFS_THREADS) return BFS_THREADS;
    const uint32_t per = (n_in + nblocks - 1) / nblocks;
    return per < 1u ? 1u : per;
}

__device__ __forceinline__ GlobalQueue out_queue(const BfsArgs& a, int ph) {
    GlobalQueue q;
    q.keys = a.qk[(ph + 1) & 1];
    q.vals = a.qv[(ph + 1) & 1];
    q.len = &a.ctr->qlen[(ph + 1) & 3];
    q.cap = a.qcap[(ph + 1) & 1];
    q.overflow = &a.ctr->overflow;
    q.coh = a.coh_out != 0;
    return q;
}

// Workgroup sum of a per-thread u64, added once to *dst.
__device__ __forceinline__ void wg_add_u64(unsigned long long v, unsigned long long* dst, unsigned long long* red) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < BFS_WAVES; ++w) t += red[w];
        if (t) atomic_add_u64(dst, t);
    }
}

// Grid-tier deferral (wave-uniform): lanes whose vertex has degree >= grid_th
// hand its adjacency list to the grid tier of phase ph and drop their own
// expansion.  One 64-bit atomic per wave allocates, in the same order, entry
// slots (high word) and HEAVY_CHUNK-edge chunk ranges (low word); each lane
// records (vertex, first chunk) and points its chunks at its entry, so the
// grid-tier kernel maps any chunk to its vertex with one load.
__device__ __forceinline__ void defer_heavy(const BfsArgs& a, int ph, int32_t v, int32_t& deg, int32_t grid_th) {
    const bool heavy = deg >= grid_th;
    const uint64_t m = __ballot(heavy);
    if (m == 0) return;
    const int lane = lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    const uint32_t nch = heavy ? (uint32_t)((deg + HEAVY_CHUNK - 1) / HEAVY_CHUNK) : 0u;
    const uint32_t ch_incl = (uint32_t)wave_inclusive_scan((int)nch);
    const uint32_t ch_total = (uint32_t)__shfl((int)ch_incl, kWave - 1, kWave);
    unsigned long long base = 0;
    if (lane == leader)
        base = __hip_atomic_fetch_add(&a.ctr->heavy_alloc[ph & 1],
                                      ((unsigned long long)__popcll(m) << 32) | (unsigned long long)ch_total,
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t base_lo = (uint32_t)wave_bcast((int)(uint32_t)base, leader);
    const uint32_t base_hi = (uint32_t)wave_bcast((int)(uint32_t)(base >> 32), leader);
    if (heavy) {
        const uint32_t entry = base_hi + (uint32_t)lane_prefix(m);
        const uint32_t c0 = base_lo + ch_incl - nch;
        if (entry < a.heavy_cap && c0 + nch <= a.chunk_cap) {
            const bool coh = a.coh_out != 0;
            coh_st2(&a.heavy[entry], make_int2(v, (int)c0), coh);
            for (uint32_t c = 0; c < nch; ++c) coh_st(&a.chunk_map[c0 + c], (int32_t)entry, coh);
        } else {
            st_relaxed(&a.ctr->overflow, 1u);
        }
        deg = 0;
    }
}

// Top-down discovery sink: exact claim, labels, optional frontier bitmap bit
// and degree accounting for the direction switch, then the LDS queue.
template <int VCAP>
struct TdSink {
    const BfsArgs& a;
    LdsQueueT<VCAP>& q;
    const GlobalQueue& out;
    uint32_t* next_bm;
    int32_t depth_next;
    bool count_edges;
    bool found;
    unsigned long long eacc;
    __device__ __forceinline__ void operator()(bool valid, int32_t v, int32_t src) {
        const bool won = valid && claim_exact(a.visited, v);
        if (won) {
            a.depth[v] = depth_next;
            a.pred[v] = src;
            found = true;
            if (next_bm) atomic_or_u32(&next_bm[v >> 5], 1u << (v & 31));
            if (count_edges) eacc += (unsigned long long)(a.offs[v + 1] - a.offs[v]);
        }
        lq_append<false>(q, out, won, v, 0);
    }
};

struct PairSink {
    LevelShared& s;
    const GlobalQueue& out;
    __device__ __forceinline__ void operator()(bool valid, int32_t nbr, int32_t src) {
        lq_append<true>(s.q, out, valid, nbr, src);
    }
};

// The deepest label: one atomic per workgroup that found a vertex.  A
// persistent launch leaves it to workgroup 0, which learns
// after the level's barrier whether the next frontier is empty (every
// vertex a top-down step claims is enqueued), so the grid's workgroups do
// not all hit one counter at every level.
template <int VCAP>
__device__ __forceinline__ void td_finish(const BfsArgs& a, int ph, TdSink<VCAP>& sink, unsigned long long* red) {
    lq_flush<false>(sink.q, sink.out, 1);
    if (sink.count_edges) wg_add_u64(sink.eacc, &a.ctr->level_edges[(ph + 1) & 3], red);
    if (a.persistent) return;
    if (__syncthreads_or(sink.found) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)sink.depth_next);
}

// Expand-contract (vertex frontier -> vertex frontier).  Neighbours are
// culled and labelled at discovery (depth = depth_next), as in
// expand_contract_atomic/cta.cuh:212-214, and compacted into the next vertex
// frontier.
// `nblocks` workgroups (the first ones of the launch) share the frontier;
// lists of at least `grid_th` edges go to the grid tier (k_heavy).
template <int VCAP>
__device__ void td_phase(const BfsArgs& a, int ph, int32_t depth_next, uint32_t* next_bm, bool count_edges,
                         LdsQueueT<VCAP>& q, GatherSmem& g, unsigned long long* red, uint32_t nblocks,
                         int32_t grid_th, uint32_t n_in_arg = kLenFromCounters) {
    const uint32_t n_in = input_len(a, ph, n_in_arg);
    phase_begin(a, ph, n_in);
    const int32_t* qin = a.qk[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    TdSink<VCAP> sink{a, q, out, next_bm, depth_next, count_edges, false, 0ull};
    const uint32_t per = entries_per_wg(a, n_in, nblocks);
    for (uint32_t base = blockIdx.x * per; base < n_in; base += nblocks * per) {
        const uint32_t i = base + threadIdx.x;
        int32_t u = -1, start = 0, deg = 0;
        if (threadIdx.x < per && i < n_in) {
            u = coh_ld(&qin[i], a.coh_in != 0);
            if (valid_vertex(a, u)) {
                start = a.offs[u];
                deg = a.offs[u + 1] - start;
            }
        }
        defer_heavy(a, ph, u, deg, grid_th);
        tiered_gather(a.cols, u, start, deg, a.cta_th, a.wave_th, g, sink);
        lq_flush<false>(q, out, LQ_CAP / 2);
    }
    td_finish(a, ph, sink, red);
}

// Contract-expand (edge frontier -> edge frontier).  Incoming (vertex,
// predecessor) pairs are culled (bitmask, history hash, exact claim) and
// labelled at contraction (depth = depth_here), then the survivors'
// neighbour lists are expanded into the outgoing edge frontier
// (contract_expand_atomic/cta.cuh:573-709).
__device__ void ce_phase(const BfsArgs& a, int ph, int32_t depth_here, LevelShared& s,
                         uint32_t n_in_arg = kLenFromCounters) {
    const uint32_t n_in = input_len(a, ph, n_in_arg);
    phase_begin(a, ph, n_in);
    const int32_t* qk = a.qk[ph & 1];
    const int32_t* qv = a.qv[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    PairSink sink{s, out};
    bool found = false;
    const uint32_t per = entries_per_wg(a, n_in, gridDim.x);
    for (uint32_t base = blockIdx.x * per; base < n_in; base += gridDim.x * per) {
        const uint32_t i = base + threadIdx.x;
        int32_t v = -1, start = 0, deg = 0;
        if (threadIdx.x < per && i < n_in) {
            v = coh_ld(&qk[i], a.coh_in != 0);
            const int32_t p = coh_ld(&qv[i], a.coh_in != 0);
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
        defer_heavy(a, ph, v, deg, a.grid_th);
        tiered_gather(a.cols, v, start, deg, a.cta_th, a.wave_th, s.g, sink);
        lq_flush<true>(s.q, out, LQ_CAP / 2);
    }
    lq_flush<true>(s.q, out, 1);
    if (__syncthreads_or(found) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)depth_here);
}

// Two-phase contract (edge frontier -> vertex frontier + labels),
// two_phase/contract_atomic/cta.cuh:494-572.
__device__ void tp_contract_phase(const BfsArgs& a, int ph, int32_t depth_here, LevelShared& s,
                                  uint32_t n_in_arg = kLenFromCounters) {
    const uint32_t n_in = input_len(a, ph, n_in_arg);
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
            v = coh_ld(&qk[i], a.coh_in != 0);
            const int32_t p = coh_ld(&qv[i], a.coh_in != 0);
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
        lq_append<false>(s.q, out, keep, v, 0);
        lq_flush<false>(s.q, out, LQ_CAP / 2);
    }
    lq_flush<false>(s.q, out, 1);
    if (__syncthreads_or(found) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)depth_here);
}

// Positional expansion of one tile of lists (workgroup-uniform call): every
// edge is output, so the tile's output size is the sum of its degrees.  One
// atomic reserves it; each position is mapped back to its list by binary
// search over the tile's ranks (staged in LDS) and written in place as
// (neighbour, source), EXPAND_ITEMS positions per thread with their loads
// issued together.  Work per workgroup follows edges, not lists.
constexpr int EXPAND_ITEMS = 8;
__device__ __forceinline__ void expand_tile_in_place(const BfsArgs& a, const GlobalQueue& out, int32_t u,
                                                     int32_t start, int32_t deg, GatherSmem& g) {
    const int tid = threadIdx.x;
    int total;
    const int rank = wg_exclusive_scan<BFS_THREADS>(deg, total, g.s_wave);
    if (total == 0) return;
    int32_t* s_rank = g.goff;
    int32_t* s_start = g.goff + BFS_THREADS;
    int32_t* s_src = g.gsrc;
    s_rank[tid] = rank;
    s_start[tid] = start;
    s_src[tid] = u;
    if (tid == 0) {
        const uint32_t base = atomic_add_u32(out.len, (uint32_t)total);
        if ((uint64_t)base + (uint64_t)total > out.cap) st_relaxed(out.overflow, 1u);
        g.o_start = (int32_t)base;
    }
    __syncthreads();
    const uint32_t base = (uint32_t)g.o_start;
    int lo = 0;
    for (int p0 = 0; p0 < total; p0 += BFS_THREADS * EXPAND_ITEMS) {
        int32_t col[EXPAND_ITEMS], src[EXPAND_ITEMS];
#pragma unroll
        for (int r = 0; r < EXPAND_ITEMS; ++r) {
            const int p = p0 + r * BFS_THREADS + tid;
            col[r] = -1;
            if (p < total) {
                // Last list with rank <= p; positions rise with r and p0.
                int hi = BFS_THREADS;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_rank[mid] <= p) lo = mid;
                    else hi = mid;
                }
                col[r] = s_start[lo] + (p - s_rank[lo]);
                src[r] = s_src[lo];
            }
        }
        int32_t v[EXPAND_ITEMS];
#pragma unroll
        for (int r = 0; r < EXPAND_ITEMS; ++r) v[r] = col[r] >= 0 ? a.cols[col[r]] : 0;
#pragma unroll
        for (int r = 0; r < EXPAND_ITEMS; ++r) {
            if (col[r] < 0) continue;
            const uint64_t pos = (uint64_t)base + (uint32_t)(p0 + r * BFS_THREADS + tid);
            if (pos < out.cap) {
                coh_st(&out.keys[pos], v[r], out.coh);
                coh_st(&out.vals[pos], src[r], out.coh);
            }
        }
    }
    __syncthreads();
}

// Two-phase expand (vertex frontier -> edge frontier of (neighbour, pred)),
// two_phase/expand_atomic/cta.cuh:606-719.  No culling: lists below the
// grid threshold are expanded in place per tile, longer ones by k_heavy.
__device__ void tp_expand_phase(const BfsArgs& a, int ph, LevelShared& s, uint32_t n_in_arg = kLenFromCounters) {
    const uint32_t n_in = input_len(a, ph, n_in_arg);
    phase_begin(a, ph, n_in);
    const int32_t* qin = a.qk[ph & 1];
    const GlobalQueue out = out_queue(a, ph);
    for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_in; base += gridDim.x * BFS_THREADS) {
        const uint32_t i = base + threadIdx.x;
        int32_t u = -1, start = 0, deg = 0;
        if (i < n_in) {
            u = coh_ld(&qin[i], a.coh_in != 0);
            if (valid_vertex(a, u)) {
                start = a.offs[u];
                deg = a.offs[u + 1] - start;
            }
        }
        defer_heavy(a, ph, u, deg, a.grid_th);
        expand_tile_in_place(a, out, u, start, deg, s.g);
    }
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
            v = coh_ld(&qk[i], a.coh_in != 0);
            p = coh_ld(&qv[i], a.coh_in != 0);
            keep = valid_vertex(a, v);
            if (keep) keep = !visited_test(a.visited, v);
            if (keep) keep = !history_seen(s.hist, v);
        }
        lq_append<true>(s.q, out, keep, v, p);
        lq_flush<true>(s.q, out, LQ_CAP / 2);
    }
    lq_flush<true>(s.q, out, 1);
}

// Bottom-up step of the direction-optimizing traversal (no counterpart in
// the reference): every unvisited vertex looks for a neighbour in the current
// frontier bitmap and stops at the first hit.  Only valid on symmetric
// graphs.  The step is a chain of dependent loads (visited word -> offsets
// -> column -> bitmap word), so latency, not bandwidth, bounds it; every
// wave therefore works on its own 64-vertex tiles with no workgroup
// barriers, so a stalled wave never holds the others:
//  * a tile whose 64 vertices are all visited is skipped after one load per
//    lane (isolated vertices of symmetric graphs are pre-marked visited by
//    the reset, so late levels skip almost every tile);
//  * stage 1: each lane checks the first BU_FIRST neighbours of its vertex
//    with independent loads;
//  * stage 2: vertices still unresolved with neighbours left are scanned by
//    BU_GROUP-lane groups of the wave, BU_GROUP neighbours per step, lowest
//    lane wins (the first frontier neighbour in list order).
// A wave owns the two bitmap words of its 64 vertices, so visited/next words
// are written without atomics.  Discovered vertices go to the vertex queue
// (one global atomic per wave tile) so that a top-down step can follow.
// Measured on RMAT-22 (ms per search): BU_FIRST 1-3 within 1 % of each
// other, 4 about 2 % slower, 6 and 8 4-10 % slower; BU_GROUP 4 and 16 no
// better than 8.
constexpr int BU_FIRST = 2;
constexpr int BU_GROUP = 8;
constexpr int BU_WAVE_GROUPS = kWave / BU_GROUP;

struct BuWave {
    int32_t start[kWave];  // pending entry: next edge to scan
    int32_t end[kWave];
    int32_t parent[kWave];
};

// Discovered vertices are buffered per wave and written to the global queue
// with one atomic per workgroup at the end (or per wave when its buffer
// fills), not one per tile: a single queue counter hit by every wave tile
// serializes the whole step.
constexpr int BU_QBUF = 512;

struct BuShared {
    BuWave w[BFS_WAVES];
    int32_t qbuf[BFS_WAVES][BU_QBUF];
    uint32_t qn[BFS_WAVES];
    uint32_t qbase[BFS_WAVES];
    unsigned long long red[BFS_WAVES];
};

// Copies `count` buffered vertices of this wave to the global queue at `base`.
__device__ __forceinline__ void bu_copy_out(const int32_t* buf, uint32_t count, uint32_t base, int32_t* qout,
                                            uint32_t qcap, uint32_t* overflow) {
    for (uint32_t i = lane_id(); i < count; i += kWave) {
        const uint32_t pos = base + i;
        if (pos < qcap)
            qout[pos] = buf[i];
        else
            st_relaxed(overflow, 1u);
    }
}

// LDS written by some lanes of a wave and read by others after it: keep the
// compiler from reordering across this point (a wave's LDS operations
// complete in program order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of a direction-optimizing level: the top-down step's queue and gather
// staging or the bottom-up step's per-wave scratch (about 16.5 KB: eight
// workgroups of 256 threads can stay resident per CU).
struct DoTdShared {
    LdsKeyQueue qv;
    GatherSmem g;
    unsigned long long red[BFS_WAVES];
};
union DoShared {
    DoTdShared td;
    BuShared bu;
};

__device__ __forceinline__ bool bm_test(const uint32_t* bm, int32_t u) { return (bm[u >> 5] >> (u & 31)) & 1u; }

__device__ void bu_waves(const BfsArgs& a, int ph, int32_t depth_next, const uint32_t* cur, uint32_t* next,
                         BuShared& s) {
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    BuWave& W = s.w[wid];
    const int grp = lane / BU_GROUP, sub = lane % BU_GROUP;
    const int gbase = lane & ~(BU_GROUP - 1);
    int32_t* qout = a.qk[(ph + 1) & 1];
    uint32_t* qlen = &a.ctr->qlen[(ph + 1) & 3];
    const uint32_t qcap = a.qcap[(ph + 1) & 1];
    int32_t* qb = s.qbuf[wid];
    uint32_t wq = 0;  // vertices in this wave's buffer (wave-uniform)
    bool found_any = false;
    unsigned long long eacc = 0;
    const int64_t step = (int64_t)gridDim.x * BFS_WAVES * kWave;
    // The next tile's visited word and row offsets are loaded one iteration
    // ahead (whether or not the vertex turns out to be visited), so the
    // current tile's scan overlaps the next tile's first two round trips.
    int64_t vb = ((int64_t)blockIdx.x * BFS_WAVES + wid) * kWave;
    uint32_t nvw = 0xFFFFFFFFu;
    int32_t noff0 = 0, noff1 = 0;
    if (vb + lane < a.n) {
        const int32_t v = (int32_t)(vb + lane);
        nvw = a.visited[v >> 5];
        noff0 = a.offs[v];
        noff1 = a.offs[v + 1];
    }
    for (; vb < a.n; vb += step) {
        const int32_t v = (int32_t)(vb + lane);
        const uint32_t vw = nvw;
        const int32_t off0 = noff0, off1 = noff1;
        if (vb + step + lane < a.n) {
            const int32_t vn = (int32_t)(vb + step + lane);
            nvw = a.visited[vn >> 5];
            noff0 = a.offs[vn];
            noff1 = a.offs[vn + 1];
        } else {
            nvw = 0xFFFFFFFFu;  // past the end: never active
        }
        const bool active = (v < a.n) && !((vw >> (v & 31)) & 1u);
        if (__ballot(active) == 0) continue;
        int32_t start = 0, deg = 0, parent = -1;
        if (active) {
            start = off0;
            deg = off1 - off0;
        }
        // Stage 1: the first BU_FIRST neighbours, loads issued together.
        if (deg > 0) {
            int32_t u[BU_FIRST];
#pragma unroll
            for (int k = 0; k < BU_FIRST; ++k) u[k] = (k < deg) ? a.cols[start + k] : -1;
#pragma unroll
            for (int k = BU_FIRST - 1; k >= 0; --k)
                if (u[k] >= 0 && bm_test(cur, u[k])) parent = u[k];
        }
        // Stage 2: the wave's unresolved lists, one per BU_GROUP-lane group.
        const bool pend = parent < 0 && deg > BU_FIRST;
        const uint64_t pm = __ballot(pend);
        if (pm) {
            const int slot = lane_prefix(pm);
            if (pend) {
                W.start[slot] = start + BU_FIRST;
                W.end[slot] = start + deg;
            }
            wave_lds_sync();
            const int np = __popcll(pm);
            for (int i = grp; i < np; i += BU_WAVE_GROUPS) {
                const int32_t e1 = W.end[i];
                int32_t hit_u = -1;
                for (int32_t e0 = W.start[i]; e0 < e1; e0 += BU_GROUP) {
                    const int32_t e = e0 + sub;
                    const int32_t u = (e < e1) ? a.cols[e] : -1;
                    const bool hit = u >= 0 && bm_test(cur, u);
                    const uint64_t hm = (__ballot(hit) >> gbase) & ((1ull << BU_GROUP) - 1);
                    if (hm) {
                        hit_u = __shfl(u, gbase + __ffsll((unsigned long long)hm) - 1, kWave);
                        break;
                    }
                }
                if (sub == 0) W.parent[i] = hit_u;
            }
            wave_lds_sync();
            if (pend) parent = W.parent[slot];
            wave_lds_sync();  // the next tile rewrites the scratch
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
            const int64_t w = v >> 5;
            if (w < a.words) {
                const uint32_t bits = (lane == 0) ? (uint32_t)fm : (uint32_t)(fm >> 32);
                next[w] = bits;
                if (bits) a.visited[w] = vw | bits;
            }
        }
        if (fm) {
            const uint32_t nf = (uint32_t)__popcll(fm);
            if (wq + nf > (uint32_t)BU_QBUF) {  // buffer full: flush this wave's part now
                uint32_t base = 0;
                if (lane == 0) base = atomic_add_u32(qlen, wq);
                base = (uint32_t)wave_bcast((int)base, 0);
                bu_copy_out(qb, wq, base, qout, qcap, &a.ctr->overflow);
                wq = 0;
                wave_lds_sync();
            }
            if (found) qb[wq + (uint32_t)lane_prefix(fm)] = v;
            wq += nf;
        }
    }
    // One queue reservation for the whole workgroup's buffered vertices.
    if (lane == 0) s.qn[wid] = wq;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < BFS_WAVES; ++w) {
            s.qbase[w] = total;
            total += s.qn[w];
        }
        const uint32_t base = total ? atomic_add_u32(qlen, total) : 0u;
        for (int w = 0; w < BFS_WAVES; ++w) s.qbase[w] += base;
    }
    __syncthreads();
    bu_copy_out(qb, wq, s.qbase[wid], qout, qcap, &a.ctr->overflow);
    wg_add_u64(eacc, &a.ctr->level_edges[(ph + 1) & 3], s.red);
    if (__syncthreads_or(found_any) && threadIdx.x == 0) atomic_max_u32(a.max_depth, (uint32_t)depth_next);
}

// Direction for level ph of the direction-optimizing traversal (Beamer,
// Asanovic, Patterson, SC 2012).  Every workgroup evaluates the same counters
// (finalized by the previous launch), so all agree without host involvement.
// m_f = degree sum of the frontier, m_u = edges of unvisited vertices;
// top-down -> bottom-up when m_f > m_u / alpha, back when n_f < n / beta.
__device__ __forceinline__ uint32_t do_direction(const BfsArgs& a, int ph, uint32_t n_f) {
    if (!a.symmetric) return 0;
    const long long m_f = (long long)a.ctr->level_edges[ph & 3];
    const long long vis = (long long)a.ctr->vis_edges[ph & 3];
    const uint32_t prev = a.ctr->dir[(ph + 3) & 3];
    long long m_u = a.edges - vis - m_f;
    if (m_u < 0) m_u = 0;
    if (prev == 0) return (m_f * (long long)a.do_alpha > m_u) ? 1u : 0u;
    return ((long long)n_f * (long long)a.do_beta < (long long)a.n) ? 0u : 1u;
}

// ===========================================================================
// Kernel bodies.
// ===========================================================================

enum SeedMode { SEED_VERTEX = 0, SEED_EDGE = 1, SEED_DO = 2 };

// Reset's label initialisation fused with the search prologue (the
// reference's MemsetKernel x 2, csr_problem.cuh:723-745, then the
// iteration-0 prologue of contract_expand_atomic/kernel.cuh:307-332 and
// expand_contract_atomic/kernel.cuh:305-319) in one launch: depth and
// predecessor := -1, visited mask and the first two frontier bitmaps := 0,
// except the source, which vertex modes label at once (depth 0, pred -2,
// visited, and its frontier bit for the direction-optimizing mode); edge
// mode leaves it unlabelled and enqueues (src, -2) for contraction.  Every
// value is written by exactly one thread, so no ordering is needed inside
// the launch.  Workgroup 0 also resets the counters; the grid barrier's
// release epoch is left alone (it grows over the problem's lifetime).
__device__ void reset_seed_body(const BfsArgs& a, int ph, int32_t src, int mode) {
    const bool label_src = mode != SEED_EDGE;
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        const bool s = label_src && i == src;
        a.depth[i] = s ? 0 : -1;
        a.pred[i] = s ? -2 : -1;
    }
    const uint32_t sbit = label_src ? (1u << (src & 31)) : 0u;
    const int32_t sw = src >> 5;
    uint32_t* fcur = a.fbm[ph % 3];
    uint32_t* fnext = a.fbm[(ph + 1) % 3];
    for (int32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < a.words; w += stride) {
        // Isolated vertices of a symmetric graph can never be reached:
        // marking them visited lets bottom-up steps skip them (their labels
        // stay -1).  The source keeps only its own seeding.
        uint32_t vis = a.iso ? a.iso[w] : 0u;
        if (w == sw) vis = (vis & ~(1u << (src & 31))) | sbit;
        a.visited[w] = vis;
        fcur[w] = (mode == SEED_DO && w == sw) ? sbit : 0u;
        fnext[w] = 0u;
    }
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DevCounters* c = a.ctr;
    for (int i = 0; i < 4; ++i) {
        c->qlen[i] = 0;
        c->level_edges[i] = 0;
        c->level_verts[i] = 0;
        c->vis_edges[i] = 0;
        c->dir[i] = 0;
    }
    c->heavy_alloc[0] = 0;
    c->heavy_alloc[1] = 0;
    c->ph_base = 0;
    c->overflow = 0;
    c->last_level = 0;
    c->levels_run = 0;
    c->saturated = 0;
    c->queued = 0;
    c->next_edges = 0;
    c->next_edges_prev = 0;
    c->bar_abort = 0;
    c->corrupt = 0;
    *a.max_depth = 0;
    a.qk[ph & 1][0] = src;
    if (mode == SEED_EDGE)
        a.qv[ph & 1][0] = -2;
    else
        c->level_edges[ph & 3] = (unsigned long long)(a.offs[src + 1] - a.offs[src]);
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

// In a captured, replayed level loop (dev_base set) a kernel's phase
// argument is relative to ctr->ph_base, which the loop's last node advances,
// and its depth argument is relative to that phase.
__device__ __forceinline__ void resolve_phase(const BfsArgs& a, int& ph, int32_t& depth) {
    if (a.dev_base) {
        ph += (int)ld_relaxed(&a.ctr->ph_base);
        depth += ph;
    }
}

enum PhaseKind { PK_TD = 0, PK_CE = 2, PK_TP_CONTRACT = 3, PK_TP_EXPAND = 4, PK_TP_FILTER = 5 };

template <int KIND>
__device__ void phase_body(BfsArgs a, int ph, int32_t depth, int slot, LevelShared& s) {
    resolve_phase(a, ph, depth);
    unsigned long long t0 = 0;
    stats_begin(t0, a);
    shared_init(s);
    if (KIND == PK_TD) td_phase(a, ph, depth, nullptr, false, s.q, s.g, s.red, gridDim.x, a.grid_th);
    if (KIND == PK_CE) ce_phase(a, ph, depth, s);
    if (KIND == PK_TP_CONTRACT) tp_contract_phase(a, ph, depth, s);
    if (KIND == PK_TP_EXPAND) tp_expand_phase(a, ph, s);
    if (KIND == PK_TP_FILTER) tp_filter_phase(a, ph, s);
    stats_end(t0, a, slot);
}

// One level of the direction-optimizing traversal: the workgroups agree on
// the direction from device counters, clear the bitmap two levels ahead and
// run a top-down or bottom-up step.  An empty frontier makes it a no-op, so
// the host can launch levels ahead of the termination check.
__device__ void do_level_body(BfsArgs a, int ph, int slot, int grid_tier, DoShared& s) {
    int32_t unused_depth = 0;
    resolve_phase(a, ph, unused_depth);
    const uint32_t n_f = phase_input_len(a, ph);
    if (n_f == 0) {
        // Keep the slot rotation going so later no-op levels also read 0.
        phase_begin(a, ph, 0);
        return;
    }
    unsigned long long t0 = 0;
    stats_begin(t0, a);
    const uint32_t dir = do_direction(a, ph, n_f);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctr->dir[ph & 3] = dir;
        a.ctr->vis_edges[(ph + 1) & 3] = a.ctr->vis_edges[ph & 3] + a.ctr->level_edges[ph & 3];
    }
    uint32_t* clear = a.fbm[(ph + 2) % 3];
    for (int32_t w = blockIdx.x * BFS_THREADS + threadIdx.x; w < a.words; w += gridDim.x * BFS_THREADS) clear[w] = 0;
    if (dir == 0) {
        // Top-down steps use the first td_grid workgroups; the launch is
        // sized for the bottom-up step, which needs every wave slot.
        const uint32_t tdg = min(gridDim.x, (uint32_t)max(1, a.td_grid));
        if (blockIdx.x < tdg) {
            if (threadIdx.x == 0) s.td.qv.count = 0;
            gather_init(s.td.g);
            __syncthreads();
            td_phase(a, ph, ph + 1, a.fbm[(ph + 1) % 3], true, s.td.qv, s.td.g, s.td.red, tdg,
                     grid_tier ? a.grid_th : INT_MAX);
        }
    } else {
        phase_begin(a, ph, 0);
        bu_waves(a, ph, ph + 1, a.fbm[ph % 3], a.fbm[(ph + 1) % 3], s.bu);
    }
    stats_end(t0, a, slot);
}

// Grid tier: the adjacency lists deferred by phase ph, cut into
// HEAVY_CHUNK-edge chunks; workgroups take chunks grid-stride and find each
// chunk's vertex through the chunk map.  Every chunk is checked against the
// entry it names, so stale map words can never index outside a list.
enum HeavyKind { HK_TD = 0, HK_TD_DO = 1, HK_PAIRS = 2 };

// Top-down grid-tier launches need only a vertex queue; a small LDS block
// keeps them at full residency.
struct TdHeavyShared {
    LdsKeyQueue qv;
    unsigned long long red[BFS_WAVES];
};
template <int KIND>
struct HeavySharedOf {
    typedef TdHeavyShared type;
};
// Pair grid tier: a batch of the workgroup's chunks is measured first, then
// reserved in the output queue with one atomic and written in place.
constexpr int HEAVY_BATCH = 1024;  // chunks per reservation
struct PairsHeavyShared {
    int32_t u[HEAVY_BATCH];
    int32_t start[HEAVY_BATCH];
    uint32_t off[HEAVY_BATCH + 1];  // exclusive offsets of the chunks' lengths; [n] = total
    int s_wave[BFS_WAVES];
    uint32_t base;
};
template <>
struct HeavySharedOf<HK_PAIRS> {
    typedef PairsHeavyShared type;
};

template <int KIND>
__device__ void heavy_body(BfsArgs a, int ph, int32_t depth_next, int slot, typename HeavySharedOf<KIND>::type& s) {
    resolve_phase(a, ph, depth_next);
    const unsigned long long alloc =
        __hip_atomic_load(&a.ctr->heavy_alloc[ph & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t H = min((uint32_t)(alloc >> 32), a.heavy_cap);
    const uint32_t C = min((uint32_t)alloc, a.chunk_cap);
    if (H == 0 || C == 0) return;
    unsigned long long t0 = 0;
    stats_begin(t0, a);
    if constexpr (KIND != HK_PAIRS) {
        if (threadIdx.x == 0) s.qv.count = 0;
        __syncthreads();
    }
    const GlobalQueue out = out_queue(a, ph);
    // HEAVY_ROUND chunks per iteration (grid-strided), each load stage issued
    // for all of them before the next, so that many dependent chains
    // (chunk map -> entry -> offsets -> columns) are in flight per workgroup.
    constexpr int R = HEAVY_ROUND;
    const uint32_t stride = gridDim.x;
    auto chunks = [&](uint32_t c0, int32_t (&u)[R], int32_t (&v)[R], bool (&valid)[R]) {
        int32_t entry[R], start[R], len[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t c = c0 + (uint32_t)r * stride;
            entry[r] = c < C ? coh_ld(&a.chunk_map[c], a.coh_out != 0) : -1;
        }
        int2 eu[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            eu[r] = (uint32_t)entry[r] < H ? coh_ld2(&a.heavy[entry[r]], a.coh_out != 0) : make_int2(-1, 0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t c = c0 + (uint32_t)r * stride;
            u[r] = -1;
            start[r] = 0;
            len[r] = 0;
            if ((uint32_t)entry[r] < H && valid_vertex(a, eu[r].x) && (uint32_t)eu[r].y <= c) {
                const int32_t s0 = a.offs[eu[r].x], deg = a.offs[eu[r].x + 1] - s0;
                const long long off = (long long)(c - (uint32_t)eu[r].y) * HEAVY_CHUNK;
                if (off < deg) {
                    u[r] = eu[r].x;
                    start[r] = s0 + (int32_t)off;
                    len[r] = min((int32_t)HEAVY_CHUNK, deg - (int32_t)off);
                }
            }
        }
        const int32_t k = (int32_t)threadIdx.x;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            valid[r] = k < len[r];
            v[r] = valid[r] ? a.cols[start[r] + k] : 0;
        }
    };
    if constexpr (KIND == HK_PAIRS) {
        // Every edge of a pair chunk is output, so a chunk's output size is
        // its length.  Per batch of HEAVY_BATCH of this workgroup's chunks
        // (c = blockIdx.x + i * stride): each thread measures the chunks
        // i = 4*tid .. 4*tid+3, a workgroup scan gives every chunk its offset,
        // one atomic reserves the batch, and the edges are written in place
        // (consecutive lanes, consecutive slots), HEAVY_ROUND chunks at a time.
        (void)chunks;
        const int tid = (int)threadIdx.x;
        // chunks of this workgroup: c = blockIdx.x + j * stride < C
        const uint32_t mine = blockIdx.x < C ? (C - blockIdx.x + stride - 1) / stride : 0u;
        constexpr int PER = HEAVY_BATCH / BFS_THREADS;
        for (uint32_t b0 = 0; b0 < mine; b0 += HEAVY_BATCH) {
            const uint32_t nb = min((uint32_t)HEAVY_BATCH, mine - b0);
            int32_t lens[PER];
            int32_t mysum = 0;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const uint32_t i = (uint32_t)(tid * PER + q);
                int32_t u = -1, start = 0, len = 0;
                if (i < nb) {
                    const uint32_t c = blockIdx.x + (b0 + i) * stride;
                    const int32_t entry = coh_ld(&a.chunk_map[c], a.coh_out != 0);
                    if ((uint32_t)entry < H) {
                        const int2 eu = coh_ld2(&a.heavy[entry], a.coh_out != 0);
                        if (valid_vertex(a, eu.x) && (uint32_t)eu.y <= c) {
                            const int32_t s0 = a.offs[eu.x], deg = a.offs[eu.x + 1] - s0;
                            const long long off = (long long)(c - (uint32_t)eu.y) * HEAVY_CHUNK;
                            if (off < deg) {
                                u = eu.x;
                                start = s0 + (int32_t)off;
                                len = min((int32_t)HEAVY_CHUNK, deg - (int32_t)off);
                            }
                        }
                    }
                    s.u[i] = u;
                    s.start[i] = start;
                }
                lens[q] = len;
                mysum += len;
            }
            int total;
            int run = wg_exclusive_scan<BFS_THREADS>(mysum, total, s.s_wave);
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const uint32_t i = (uint32_t)(tid * PER + q);
                if (i < nb) s.off[i] = (uint32_t)run;
                run += lens[q];
            }
            if (tid == 0) {
                s.off[nb] = (uint32_t)total;
                s.base = total ? atomic_add_u32(out.len, (uint32_t)total) : 0u;
                if (total && (uint64_t)s.base + (uint64_t)total > out.cap) st_relaxed(out.overflow, 1u);
            }
            __syncthreads();
            const uint32_t base = s.base;
            for (uint32_t i0 = 0; i0 < nb; i0 += R) {
                int32_t v[R];
                bool valid[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t i = i0 + r;
                    const int32_t len = i < nb ? (int32_t)(s.off[i + 1] - s.off[i]) : 0;
                    valid[r] = tid < len;
                    v[r] = valid[r] ? a.cols[s.start[i] + tid] : 0;
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (!valid[r]) continue;
                    const uint32_t i = i0 + r;
                    const uint64_t pos = (uint64_t)base + s.off[i] + (uint32_t)tid;
                    if (pos < out.cap) {
                        coh_st(&out.keys[pos], v[r], out.coh);
                        coh_st(&out.vals[pos], s.u[i], out.coh);
                    }
                }
            }
            __syncthreads();
        }
    } else {
        TdSink<1> tds{a, s.qv, out, (KIND == HK_TD_DO) ? a.fbm[(ph + 1) % 3] : nullptr, depth_next, KIND == HK_TD_DO,
                   false, 0ull};
        for (uint32_t c0 = blockIdx.x; c0 < C; c0 += stride * R) {
            int32_t u[R], v[R];
            bool valid[R];
            chunks(c0, u, v, valid);
#pragma unroll
            for (int r = 0; r < R; ++r) tds(valid[r], v[r], u[r]);
            lq_flush<false>(s.qv, out, LQ_CAP / 2);
        }
        td_finish(a, ph, tds, s.red);
    }
    stats_end(t0, a, slot);
}

// Persistent kernels: all levels in one launch of a co-resident grid with a software
// grid barrier between phases (contract_expand_atomic::KernelGlobalBarrier,
// kernel.cuh:281-554; two_phase::Kernel, two_phase/kernel.cuh:60-313).
// The loop control (the input frontier's length) is read by one thread per
// workgroup and shared through LDS after each barrier; every workgroup reads
// the same value, so all leave the loop together.  With `sat_limit` set, the
// kernel quits when a frontier exceeds it (SATURATION_QUIT) and records
// where it stopped.
//
// Grid tier: a level's expansion defers lists of at least grid_th edges (as
// the level kernels do); after the level's barrier every workgroup reads the
// same deferral count, and when it is nonzero the whole grid splits those
// lists into chunks (heavy_body) before one more barrier.  Levels without
// hubs pay no extra barrier.  grid_th = INT_MAX turns the tier off.
enum FusedKind { FK_TD = 0, FK_CE = 1, FK_TP = 2 };

template <int KIND>
struct FusedHeavy {
    static constexpr int kind = (KIND == FK_TD) ? HK_TD : HK_PAIRS;
    typedef typename HeavySharedOf<kind>::type shared;
};

template <int KIND>
__device__ void fused_body(const BfsArgs& a, int ph, int32_t depth, int max_levels, uint32_t sat_limit, LevelShared& s,
                           typename FusedHeavy<KIND>::shared& hs) {
    unsigned long long t0 = 0, busy = 0;
    if (a.wg_stats && threadIdx.x == 0) t0 = wg_clock();
    shared_init(s);
    // Every phase passes its frontier through agent-scope accesses (the
    // host sets a.persistent, a.coh_in and a.coh_out) and ends with the light
    // barrier.  `a` is read-only here: a kernel that writes into its argument
    // block gets a private copy of all of it, and every field read becomes a
    // scratch load (that cost each persistent strategy a quarter of its rate).
    const GridBarrier bar{a.bar_arrive, &a.ctr->bar_release, &a.ctr->bar_abort, 1u};
    uint32_t epoch = barrier_base_epoch(bar);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctr->saturated = 0;
    // Barrier payload: the length of the frontier queue in counter slot
    // `slot` (clamped to 31 bits: a frontier never holds 2^31 entries, since
    // E < 2^31) with bit 31 set when the grid tier has lists deferred by
    // phase `heavy_ph` (heavy_ph < 0: not asked).  Workgroup 0 reads both
    // after every workgroup has arrived.
    auto payload_for = [&a](int slot, int heavy_ph) {
        return [&a, slot, heavy_ph]() -> uint32_t {
            uint32_t len = ld_relaxed(&a.ctr->qlen[slot & 3]);
            if (len > 0x7FFFFFFFu) len = 0x7FFFFFFFu;
            const bool deferred = heavy_ph >= 0 && __hip_atomic_load(&a.ctr->heavy_alloc[heavy_ph & 1], __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT) != 0ull;
            return len | (deferred ? 0x80000000u : 0u);
        };
    };
    const bool grid_tier = a.grid_th != INT_MAX;
    // Every thread reads the first level's length itself (written by the
    // launch before this one); later lengths come with the barriers.
    uint32_t n_in = ld_relaxed(&a.ctr->qlen[ph & 3]);
    int level = 0;
    for (; level < max_levels; ++level) {
        if (n_in == 0 || n_in > a.qcap[ph & 1]) break;
        if (sat_limit && n_in > sat_limit) {
            if (blockIdx.x == 0 && threadIdx.x == 0) a.ctr->saturated = 1;
            break;
        }
        unsigned long long tb = 0;
        if (a.wg_stats && threadIdx.x == 0) tb = wg_clock();
        int heavy_ph = ph;          // phase whose deferred lists the grid tier takes
        int32_t heavy_depth = depth;  // depth label of what it discovers (top-down)
        uint32_t word = 0;
        if (KIND == FK_TD) {
            td_phase(a, ph, depth + 1, nullptr, false, s.q, s.g, s.red, gridDim.x, a.grid_th, n_in);
            heavy_depth = depth + 1;
            ph += 1;
        } else if (KIND == FK_CE) {
            ce_phase(a, ph, depth, s, n_in);
            ph += 1;
        } else {
            tp_contract_phase(a, ph, depth, s, n_in);
            if (a.wg_stats && threadIdx.x == 0) busy += wg_clock() - tb;
            if (!grid_sync(bar, ++epoch, payload_for(ph + 1, -1), &s.ctl, word)) break;
            if (a.wg_stats && threadIdx.x == 0) tb = wg_clock();
            // An overflowed vertex queue counts as empty (as phase_input_len
            // has it); the overflow flag ends the search.
            uint32_t n_mid = word & 0x7FFFFFFFu;
            if (n_mid > a.qcap[(ph + 1) & 1]) n_mid = 0;
            tp_expand_phase(a, ph + 1, s, n_mid);
            heavy_ph = ph + 1;
            ph += 2;
        }
        depth += 1;
        if (a.wg_stats && threadIdx.x == 0) busy += wg_clock() - tb;
        if (!grid_sync(bar, ++epoch, payload_for(ph, grid_tier ? heavy_ph : -1), &s.ctl, word)) break;
        if (word & 0x80000000u) {
            // Lists this level deferred: the whole grid takes them in
            // chunks, appending to the same next frontier.
            heavy_body<FusedHeavy<KIND>::kind>(a, heavy_ph, heavy_depth, 0, hs);
            if (!grid_sync(bar, ++epoch, payload_for(ph, -1), &s.ctl, word)) break;
        }
        n_in = word & 0x7FFFFFFFu;
        // Top-down: the level claimed vertices (at `depth`) iff the next
        // frontier is not empty (td_finish leaves this to workgroup 0).
        if (KIND == FK_TD && n_in != 0 && blockIdx.x == 0 && threadIdx.x == 0)
            atomic_max_u32(a.max_depth, (uint32_t)depth);
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

// Global-barrier benchmark (test/gbarrier/test_gbarrier.cu:43-50): the
// kernel does nothing but pass `iterations` barriers.
__device__ void gbarrier_bench_body(const GridBarrier& bar, int iterations, uint32_t* s_word) {
    uint32_t epoch = barrier_base_epoch(bar);
    for (int i = 0; i < iterations; ++i)
        if (!grid_sync(bar, ++epoch, s_word)) break;
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

__device__ void gbarrier_check_body(const GridBarrier& bar, uint32_t* data, int words, int iterations,
                                    unsigned long long* mismatches, uint32_t* s_word) {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    uint32_t epoch = barrier_base_epoch(bar);
    unsigned long long bad = 0;
    const bool coh = bar.light != 0;  // light barrier: the slabs go through agent-scope accesses
    for (int it = 0; it < iterations; ++it) {
        for (int j = threadIdx.x; j < words; j += BFS_THREADS)
            coh_st((int32_t*)&data[(size_t)b * words + j], (int32_t)gb_pattern(it, b, j), coh);
        // The payload is the iteration number as workgroup 0 saw it in a
        // word every workgroup wrote before arriving: all must agree.
        uint32_t seen = 0;
        if (!grid_sync(bar, ++epoch,
                       [&] {
                           return (uint32_t)coh_ld((const int32_t*)&data[(size_t)(G - 1) * words], coh) ^
                                  gb_pattern(it, G - 1, 0) ^ (uint32_t)it;
                       },
                       s_word, seen))
            break;
        bad += (seen != (uint32_t)it) ? 1u : 0u;
        const uint32_t others[3] = {(b + 1) % G, (b + G / 2 + 1) % G, (b + G - 1) % G};
        for (int k = 0; k < 3; ++k) {
            const uint32_t o = others[k];
            for (int j = threadIdx.x; j < words; j += BFS_THREADS)
                bad += ((uint32_t)coh_ld((const int32_t*)&data[(size_t)o * words + j], coh) != gb_pattern(it, o, j))
                           ? 1u : 0u;
        }
        if (!grid_sync(bar, ++epoch, s_word)) break;
    }
    if (bad) atomic_add_u64(mismatches, bad);
}

}  // namespace b40c
