// Device-side building blocks of the BFS level kernels for gfx950 (wave64).
//
//  * LdsQueue: frontier-queue allocation.  Waves reserve slots in an LDS
//    buffer with one ballot + one LDS atomic per wave; the workgroup flushes
//    the buffer to the global queue with ONE global atomic per flush
//    (replaces CooperativeTileScan::ScanTileWithEnqueue's per-tile atomic,
//    b40c/util/scan/cooperative_scan.cuh:173-229).  A wave that finds the LDS
//    buffer full spills straight to the global queue (one atomic per wave).
//  * tiered_gather: neighbour gathering in three tiers (whole workgroup for
//    degree >= cta threshold, whole wave for degree >= wave threshold,
//    scan-ranked gather staged in LDS for the rest), replacing
//    ExpandByCta / ExpandByWarp / ExpandByScan
//    (b40c/graph/bfs/two_phase/expand_atomic/cta.cuh:215-423, 606-719).
//    Thresholds are runtime values.
//  * claim_exact: the visited bitmask with an agent-scope atomicOr, so each
//    vertex has exactly one discoverer (the reference's best-effort bitmask
//    RMW, contract_expand_atomic/cta.cuh:237-252, lets duplicates through).
//  * HistoryHash: per-workgroup LDS hash of recently seen ids that culls
//    duplicates before they reach the atomic (the reference's warp/history
//    culls, contract_expand_atomic/cta.cuh:363-387,
//    two_phase/contract_atomic/cta.cuh:276-297, sized for wave64 + LDS).
#pragma once

#include <cstddef>

#include "common.hpp"

namespace b40c {

constexpr int BFS_THREADS = 256;
constexpr int BFS_WAVES = BFS_THREADS / kWave;
constexpr int LQ_CAP = 2048;          // LDS queue entries
constexpr int GATHER_ELEMS = 1024;    // fine-tier staging window
constexpr int HIST_SIZE = 1024;       // history hash entries

struct BfsArgs {
    const int32_t* offs;
    const int32_t* cols;
    int32_t* depth;
    int32_t* pred;
    uint32_t* visited;       // 1 bit per vertex
    uint32_t* fbm[3];        // frontier bitmaps (direction-optimizing), rotating by level % 3
    int2* heavy;             // grid-tier entries (vertex, first chunk), in chunk order
    int32_t* chunk_map;      // grid-tier chunk -> entry index
    uint32_t heavy_cap;      // capacity of heavy[] (entries)
    uint32_t chunk_cap;      // capacity of chunk_map[]
    int32_t* qk[2];          // frontier queue keys (ping-pong)
    int32_t* qv[2];          // frontier queue values (predecessors), edge frontiers only
    uint32_t qcap[2];        // capacities of qk[i] (elements)
    DevCounters* ctr;
    uint32_t* bar_arrive;    // persistent kernels: grid-barrier arrival word per workgroup
    DevCounters* hctr;       // device address of the pinned host mirror of *ctr
    uint32_t* max_depth;     // deepest depth label written
    unsigned long long* wg_stats;  // instrumented: [grid][2] (busy clocks, lifetime clocks)
    int32_t n;
    int32_t words;           // ceil(n / 32)
    int32_t cta_th;
    int32_t wave_th;
    int32_t grid_th;         // degree >= this: grid tier (split across workgroups)
    int32_t symmetric;       // graph has every reverse edge (bottom-up allowed)
    int32_t do_alpha;
    int32_t do_beta;
    int32_t td_grid;         // direction-optimizing: workgroups of a top-down step
    const uint32_t* iso;     // symmetric graphs: isolated-vertex bitmap, pre-marked visited (else null)
    int32_t dev_base;        // 1: phase = ctr->ph_base + kernel argument (captured, replayed level loops)
    int32_t persistent;      // 1 inside a persistent (single-launch) search
    int32_t coh_in;          // persistent: frontier inputs are read with coh_ld
    int32_t coh_out;         // persistent: outputs (queue, grid-tier entries) are written with coh_st
    long long edges;
};

// Frontier data that other workgroups of the same persistent launch read
// after a grid barrier (queue entries, grid-tier entries and chunk maps).
// With `coh` the accesses are agent-scope atomic ones, which the GPU keeps
// coherent across XCDs without any cache maintenance, so the persistent
// kernels' barrier only has to wait for each wave's stores to complete
// (grid_sync's light mode) instead of writing back and invalidating its
// XCD's L2 once per workgroup.  Without `coh` (separate launches) they are
// plain accesses.
__device__ __forceinline__ void coh_st(int32_t* p, int32_t v, bool coh) {
    if (coh) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}
__device__ __forceinline__ int32_t coh_ld(const int32_t* p, bool coh) {
    return coh ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
}
__device__ __forceinline__ void coh_st2(int2* p, int2 v, bool coh) {
    if (coh) {
        const unsigned long long w = (unsigned long long)(uint32_t)v.x | ((unsigned long long)(uint32_t)v.y << 32);
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *p = v;
    }
}
__device__ __forceinline__ int2 coh_ld2(const int2* p, bool coh) {
    if (!coh) return *p;
    const unsigned long long w = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
    return make_int2((int32_t)(uint32_t)w, (int32_t)(uint32_t)(w >> 32));
}

constexpr int HEAVY_CHUNK = 256;     // grid-tier edges per work item (one per thread)
constexpr int HEAVY_ROUND = 4;       // grid-tier work items a workgroup has in flight at once

// ---------------------------------------------------------------------------
// LDS-buffered queue.  VCAP is the capacity of the value array: LQ_CAP for a
// queue that can carry (key, value) pairs, 1 for a keys-only queue.  A
// pair-capable queue also serves keys-only appends (VALS = false), so one
// kernel that alternates key and pair steps uses ONE queue object; no other
// field ever shares storage with `count` or `gbase`.
// ---------------------------------------------------------------------------
template <int VCAP>
struct LdsQueueT {
    int32_t keys[LQ_CAP];
    int32_t vals[VCAP];
    uint32_t count;
    uint32_t gbase;
};
using LdsKeyQueue = LdsQueueT<1>;
using LdsPairQueue = LdsQueueT<LQ_CAP>;
static_assert(offsetof(LdsKeyQueue, count) >= sizeof(int32_t) * (LQ_CAP + 1), "key queue count overlaps its arrays");
static_assert(offsetof(LdsPairQueue, count) >= sizeof(int32_t) * 2 * LQ_CAP, "pair queue count overlaps its arrays");

struct GlobalQueue {
    int32_t* keys;
    int32_t* vals;
    uint32_t* len;
    uint32_t cap;
    uint32_t* overflow;
    bool coh = false;  // entries written with agent-scope stores (persistent launches)
};

// Wave-uniform call: every lane of the wave must be active.
template <bool VALS, int VCAP>
__device__ __forceinline__ void lq_append(LdsQueueT<VCAP>& q, const GlobalQueue& g, bool want, int32_t k,
                                          int32_t v) {
    static_assert(!VALS || VCAP == LQ_CAP, "pair appends need a pair queue");
    const uint64_t m = __ballot(want);
    if (m == 0) return;
    const int lane = lane_id();
    const int cnt = __popcll(m);
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&q.count, (uint32_t)cnt);
    base = (uint32_t)wave_bcast((int)base, leader);
    const uint32_t pos = base + (uint32_t)lane_prefix(m);
    if (want && pos < (uint32_t)LQ_CAP) {
        q.keys[pos] = k;
        if (VALS) q.vals[pos] = v;
    }
    if (base + (uint32_t)cnt > (uint32_t)LQ_CAP) {
        const uint32_t spill_start = base > (uint32_t)LQ_CAP ? base : (uint32_t)LQ_CAP;
        const uint32_t nspill = base + (uint32_t)cnt - spill_start;
        uint32_t gb = 0;
        if (lane == leader) gb = atomic_add_u32(g.len, nspill);
        gb = (uint32_t)wave_bcast((int)gb, leader);
        if (want && pos >= (uint32_t)LQ_CAP) {
            const uint32_t gp = gb + (pos - spill_start);
            if (gp < g.cap) {
                coh_st(&g.keys[gp], k, g.coh);
                if (VALS) coh_st(&g.vals[gp], v, g.coh);
            } else {
                st_relaxed(g.overflow, 1u);
            }
        }
    }
}

// Workgroup-uniform call.  Flushes when the buffer holds at least `min_fill`
// entries (min_fill = 1: flush whatever is there).
template <bool VALS, int VCAP>
__device__ __forceinline__ void lq_flush(LdsQueueT<VCAP>& q, const GlobalQueue& g, uint32_t min_fill) {
    static_assert(!VALS || VCAP == LQ_CAP, "pair flushes need a pair queue");
    __syncthreads();
    uint32_t cnt = q.count;
    __syncthreads();
    if (cnt > (uint32_t)LQ_CAP) cnt = LQ_CAP;
    if (cnt < min_fill || cnt == 0) return;
    if (threadIdx.x == 0) q.gbase = atomic_add_u32(g.len, cnt);
    __syncthreads();
    const uint32_t gb = q.gbase;
    for (uint32_t j = threadIdx.x; j < cnt; j += BFS_THREADS) {
        const uint32_t gp = gb + j;
        if (gp < g.cap) {
            coh_st(&g.keys[gp], q.keys[j], g.coh);
            if (VALS) coh_st(&g.vals[gp], q.vals[j], g.coh);
        }
    }
    if (threadIdx.x == 0 && gb + cnt > g.cap) st_relaxed(g.overflow, 1u);
    __syncthreads();
    if (threadIdx.x == 0) q.count = 0;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Culling.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool visited_test(const uint32_t* visited, int32_t v) {
    return (visited[v >> 5] >> (v & 31)) & 1u;
}

// Exactly one caller per vertex returns true.
__device__ __forceinline__ bool claim_exact(uint32_t* visited, int32_t v) {
    uint32_t* w = visited + (v >> 5);
    const uint32_t bit = 1u << (v & 31);
    if (*w & bit) return false;
    return (atomic_or_u32(w, bit) & bit) == 0;
}

struct HistoryHash {
    int32_t slot[HIST_SIZE];
};

__device__ __forceinline__ void history_init(HistoryHash& h) {
    for (int i = threadIdx.x; i < HIST_SIZE; i += BFS_THREADS) h.slot[i] = -1;
}

// True when `v` was seen recently by this workgroup (a duplicate to drop).
// Racing writers may both miss: the hash only thins duplicates, the exact
// claim decides.
__device__ __forceinline__ bool history_seen(HistoryHash& h, int32_t v) {
    const uint32_t idx = ((uint32_t)v * 2654435761u) >> (32 - 10);
    if (h.slot[idx] == v) return true;
    h.slot[idx] = v;
    return false;
}

// ---------------------------------------------------------------------------
// Tiered neighbour gather.  `sink(valid, neighbour, source)` is called
// wave-uniformly (all lanes active) for every edge of every input vertex.
// `deg` is consumed (zeroed) as vertices are handled.  Workgroup-uniform call.
// ---------------------------------------------------------------------------
struct GatherSmem {
    int32_t owner;
    int32_t o_start, o_deg, o_src;
    int32_t goff[GATHER_ELEMS];
    int32_t gsrc[GATHER_ELEMS];
    int32_t s_wave[BFS_WAVES];
};

__device__ __forceinline__ void gather_init(GatherSmem& g) {
    if (threadIdx.x == 0) g.owner = -1;
}

template <typename Sink>
__device__ __forceinline__ void tiered_gather(const int32_t* __restrict__ cols, int32_t src, int32_t start,
                                              int32_t deg, int32_t cta_th, int32_t wave_th, GatherSmem& g,
                                              Sink& sink) {
    const int tid = threadIdx.x;
    const int lane = lane_id();
    // Whole-workgroup tier: threads vie for the workgroup; the winner's list
    // is strided by all threads.
    while (true) {
        if (deg >= cta_th) g.owner = tid;
        __syncthreads();
        const int owner = g.owner;
        if (owner < 0) break;
        if (tid == owner) {
            g.o_start = start;
            g.o_deg = deg;
            g.o_src = src;
            deg = 0;
        }
        __syncthreads();
        const int32_t os = g.o_start, od = g.o_deg, osrc = g.o_src;
        for (int32_t e0 = 0; e0 < od; e0 += BFS_THREADS) {
            const int32_t e = e0 + tid;
            const bool valid = e < od;
            const int32_t v = valid ? cols[os + e] : 0;
            sink(valid, v, osrc);
        }
        __syncthreads();
        if (tid == 0) g.owner = -1;
        __syncthreads();
    }
    // Whole-wave tier: lanes vie through a ballot; the wave strides the list.
    while (true) {
        const uint64_t m = __ballot(deg >= wave_th);
        if (m == 0) break;
        const int leader = __ffsll((unsigned long long)m) - 1;
        const int32_t ls = wave_bcast(start, leader);
        const int32_t ld = wave_bcast(deg, leader);
        const int32_t lsrc = wave_bcast(src, leader);
        if (lane == leader) deg = 0;
        for (int32_t e0 = 0; e0 < ld; e0 += kWave) {
            const int32_t e = e0 + lane;
            const bool valid = e < ld;
            const int32_t v = valid ? cols[ls + e] : 0;
            sink(valid, v, lsrc);
        }
    }
    // Fine tier: workgroup scan of the remaining (small) degrees, gather
    // offsets staged in LDS, then a coalesced sweep of the staged window.
    int total;
    const int rank = wg_exclusive_scan<BFS_THREADS>(deg, total, g.s_wave);
    for (int w0 = 0; w0 < total; w0 += GATHER_ELEMS) {
        const int lo = max(rank, w0);
        const int hi = min(rank + deg, w0 + GATHER_ELEMS);
        for (int j = lo; j < hi; ++j) {
            g.goff[j - w0] = start + (j - rank);
            g.gsrc[j - w0] = src;
        }
        __syncthreads();
        const int cnt = min(GATHER_ELEMS, total - w0);
        for (int k0 = 0; k0 < cnt; k0 += BFS_THREADS) {
            const int k = k0 + tid;
            const bool valid = k < cnt;
            int32_t v = 0, s = 0;
            if (valid) {
                v = cols[g.goff[k]];
                s = g.gsrc[k];
            }
            sink(valid, v, s);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Per-workgroup clock instrumentation (KernelRuntimeStats,
// b40c/util/kernel_runtime_stats.cuh:77-113): busy clocks and lifetime.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wg_clock() { return wall_clock64(); }

}  // namespace b40c
