// Shared host/device utilities for the gfx950 BFS library: status plumbing,
// wave64 primitives, agent-scope atomics and the software grid barrier.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "b40c_mi355x.h"

namespace b40c {

// ---------------------------------------------------------------------------
// Error plumbing.  Every C entry point returns a b40c_status; the message of
// the last failure on the calling thread is kept for b40c_last_error_string
// (replaces B40CPerror, b40c/util/error_utils.cuh:44-56).
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
b40c_status fail(b40c_status code, const char* fmt, ...);

#define B40C_HIP(expr)                                                                    \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return ::b40c::fail(_e == hipErrorOutOfMemory ? B40C_ERR_OOM : B40C_ERR_HIP,  \
                                "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),    \
                                __FILE__, __LINE__);                                      \
    } while (0)

#define B40C_TRY(expr)                       \
    do {                                     \
        b40c_status _s = (expr);             \
        if (_s != B40C_OK) return _s;        \
    } while (0)

// ---------------------------------------------------------------------------
// Device-side counters shared by every traversal strategy.  Queue lengths
// rotate over four slots indexed by (level & 3): level L reads slot L, appends
// into slot L+1 and clears slot L+2 (the counter protocol of CtaWorkProgress,
// b40c/util/cta_work_progress.cuh:173-212).
// ---------------------------------------------------------------------------
struct alignas(256) DevCounters {
    uint32_t qlen[4];            // rotating queue lengths
    uint32_t overflow;           // nonzero once any reservation passed capacity
    uint32_t last_level;         // deepest level that had a nonempty input frontier
    uint32_t levels_run;         // persistent kernels: levels executed
    uint32_t saturated;          // persistent kernels: quit because the frontier got wide
    unsigned long long queued;   // total elements taken from input frontiers
    unsigned long long next_edges;  // degree sum of the vertices discovered this level
    unsigned long long next_edges_prev;
    // Grid barrier release word: low 32 bits the epoch of the last completed
    // barrier (never reset by a search), high 32 bits the word workgroup 0
    // published with it.
    unsigned long long bar_release;
    uint32_t bar_abort;          // set when a barrier wait timed out
    uint32_t corrupt;            // set when a frontier held an id outside [0, n)
    unsigned long long level_edges[4];  // rotating per-level degree sums (direction switch)
    uint32_t level_verts[4];            // rotating per-level discovered-vertex counts
    unsigned long long vis_edges[4];    // degree sum of vertices visited before the level
    uint32_t dir[4];                    // direction chosen per level (0 top-down, 1 bottom-up)
    uint32_t ph_base;                   // graph-replayed level loops: phase of the replay's first step
    uint32_t max_depth;                 // deepest depth label written by the search
    unsigned long long heavy_alloc[2];  // grid tier by level parity: (entries << 32) | chunks
};

// ---------------------------------------------------------------------------
// Agent-scope atomics (visible across all XCDs of the device).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ld_acquire(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add_u32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atomic_add_u64(unsigned long long* p, unsigned long long v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_or_u32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_max_u32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Wave64 primitives.
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int lane_prefix(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

__device__ __forceinline__ int wave_bcast(int v, int src_lane) { return __shfl(v, src_lane, kWave); }

// Inclusive prefix sum across the 64 lanes of the wave.
__device__ __forceinline__ int wave_inclusive_scan(int x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    return x;
}

// Exclusive prefix sum across the workgroup (THREADS threads).  `s_wave`
// needs THREADS/64 ints of LDS.  Contains two workgroup barriers.
template <int THREADS>
__device__ __forceinline__ int wg_exclusive_scan(int x, int& total, int* s_wave) {
    constexpr int NW = THREADS / kWave;
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    int incl = wave_inclusive_scan(x);
    if (lane == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        int t = s_wave[w];
        base += (w < wid) ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return base + incl - x;
}

// ---------------------------------------------------------------------------
// Software grid barrier for grids whose workgroups are all co-resident
// (replaces GlobalBarrier, b40c/util/global_barrier.cuh:47-127).
//
// Each workgroup except 0 owns one arrival word and stores the barrier's
// epoch there with release semantics, then polls the release word.
// Workgroup 0 polls every arrival word at once (each of its threads loads a
// strided share of them per poll, all loads in flight together), then stores
// the epoch into the release word.  No two workgroups ever update the same
// word, so there is no same-address atomic traffic; a barrier costs about
// three agent-scope round trips (arrival store, master poll, release poll).
//
// The release word is 64 bits: the epoch in the low half and, in the high
// half, a word workgroup 0 reads after every workgroup has arrived (the next
// frontier's length, say).  Waiters get it with the same load that ends
// their wait, so every workgroup starts the next phase with the same value
// and without another memory round trip.
//
// Epochs grow monotonically over the lifetime of the barrier memory: a
// kernel reads the release word's epoch once at its start (every workgroup
// reads the same value, since no barrier can complete before all have
// arrived) and passes base + 1, base + 2, ... to successive barriers.
// Arrival words therefore only ever hold epochs up to the current one, and
// equality means "arrived".  Agent scope makes the level's writes visible
// across all XCDs' L2s.
//
// The wait is bounded: after kBarrierSpinLimit polls a workgroup raises
// `abort_flag` and leaves, and waiters leave as soon as they see the flag, so
// a broken launch (for example a grid that is not co-resident) ends instead
// of hanging.  grid_sync returns false in every thread of every workgroup
// that saw the abort; callers stop their level loop then.  After an abort
// the epochs of the workgroups disagree, so the host clears the barrier
// memory before its next persistent launch.
// ---------------------------------------------------------------------------
constexpr uint32_t kBarrierSpinLimit = 1u << 22;

struct GridBarrier {
    uint32_t* arrive;              // [grid] arrival epoch of each workgroup (word 0 unused)
    unsigned long long* release;   // (payload << 32) | epoch of the last completed barrier
    uint32_t* abort_flag;          // nonzero once a wait ran out of polls
    // Light mode: every piece of data the grid passes between workgroups
    // across this barrier is written and read with agent-scope atomic
    // accesses, which stay coherent across XCDs without cache maintenance.
    // Arrival then only waits for each wave's stores to complete; the
    // agent-scope release and acquire (an L2 writeback and an L2
    // invalidate in every workgroup, serialised per XCD) are left out.
    uint32_t light;
};

__device__ __forceinline__ uint32_t barrier_base_epoch(const GridBarrier& b) {
    return (uint32_t)__hip_atomic_load(b.release, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// make_payload() runs in thread 0 of workgroup 0 after every workgroup has
// arrived (and after an acquire fence); its value reaches every thread of
// every workgroup in `payload`.  s_word: one LDS word of the calling
// workgroup, used to broadcast the payload.
template <typename PayloadFn>
__device__ __forceinline__ bool grid_sync(const GridBarrier& b, uint32_t epoch, PayloadFn make_payload,
                                          uint32_t* s_word, uint32_t& payload) {
    const bool light = b.light != 0;
    if (light) {
        // Every wave's agent-scope stores have completed (and the compiler
        // moves no memory access across this point).
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    __syncthreads();
    bool aborted = false;
    const uint32_t nb = gridDim.x;
    if (blockIdx.x == 0) {
        uint32_t spins = 0;
        while (true) {
            bool all = true;
            for (uint32_t w = 1 + threadIdx.x; w < nb; w += blockDim.x) all &= ld_relaxed(&b.arrive[w]) == epoch;
            if (all) break;
            if (ld_relaxed(b.abort_flag) || ++spins > kBarrierSpinLimit) {
                st_relaxed(b.abort_flag, 1u);
                aborted = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!light) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        aborted = __syncthreads_or(aborted) != 0;
        if (threadIdx.x == 0 && !aborted) {
            const uint32_t p = make_payload();
            *s_word = p;
            const unsigned long long rw = ((unsigned long long)p << 32) | epoch;
            if (light) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __hip_atomic_store(b.release, rw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_store(b.release, rw, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    } else if (threadIdx.x == 0) {
        if (light)
            __hip_atomic_store(&b.arrive[blockIdx.x], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            __hip_atomic_store(&b.arrive[blockIdx.x], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spins = 0;
        unsigned long long r;
        while ((uint32_t)(r = __hip_atomic_load(b.release, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != epoch) {
            if (ld_relaxed(b.abort_flag) || ++spins > kBarrierSpinLimit) {
                st_relaxed(b.abort_flag, 1u);
                aborted = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!light) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_word = (uint32_t)(r >> 32);
    }
    if (light) __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const bool ok = __syncthreads_or(aborted) == 0;
    payload = *s_word;
    __syncthreads();  // s_word may be rewritten by the next barrier
    return ok;
}

// Barrier without a payload.
__device__ __forceinline__ bool grid_sync(const GridBarrier& b, uint32_t epoch, uint32_t* s_word) {
    uint32_t unused = 0;
    return grid_sync(b, epoch, [] { return 0u; }, s_word, unused);
}

// ---------------------------------------------------------------------------
// Host helpers.
// ---------------------------------------------------------------------------
inline int ceil_log2_u64(uint64_t n) {
    int b = 0;
    while ((1ull << b) < n) ++b;
    return b;
}

struct DeviceInfo {
    int device = 0;
    int num_cus = 256;
    int arch_major = 9;
    std::string arch;
};
b40c_status device_info(DeviceInfo& info);

// Fill helpers that are safe under stream capture (kernels, not memsets).
b40c_status fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t stream);

}  // namespace b40c
