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
    uint32_t bar_count;          // grid barrier arrival counter
    uint32_t bar_gen;            // grid barrier generation
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
// Software grid barrier for cooperatively launched (co-resident) grids
// (replaces GlobalBarrier, b40c/util/global_barrier.cuh:47-127).  A counter
// plus a generation word: the last workgroup to arrive resets the counter and
// bumps the generation with release semantics; the others spin with acquire
// loads.  Agent scope makes the level's writes visible across all XCDs' L2s.
// ---------------------------------------------------------------------------
// The wait is bounded: after kBarrierSpinLimit polls a workgroup raises
// `abort_flag` and leaves, and waiters leave as soon as they see the flag, so
// a broken launch (for example a grid that is not co-resident) ends instead
// of hanging; callers stop their level loop when the flag is set.
constexpr uint32_t kBarrierSpinLimit = 1u << 24;

__device__ __forceinline__ void grid_barrier(uint32_t* count, uint32_t* gen, uint32_t nblocks, uint32_t* abort_flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = ld_relaxed(gen);
        const uint32_t arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            st_relaxed(count, 0u);
            __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            // Poll with relaxed loads (no cache invalidation per poll), then
            // one acquire fence once the generation has moved.
            uint32_t spins = 0;
            while (ld_relaxed(gen) == g) {
                if (ld_relaxed(abort_flag)) break;
                if (++spins > kBarrierSpinLimit) {
                    st_relaxed(abort_flag, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
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
