// Single-GPU BFS for gfx950: the __global__ kernels (each a thin wrapper that
// declares its LDS block and calls its body in bfs_kernels.hpp), then the
// CsrProblem object, the enactors and the C ABI (bfs_host.inc).
//
// Reference counterparts (b40c):
//   CsrProblem                b40c/graph/bfs/csr_problem.cuh
//   EnactorExpandContract     b40c/graph/bfs/enactor_expand_contract.cuh
//   EnactorContractExpand     b40c/graph/bfs/enactor_contract_expand.cuh
//   EnactorTwoPhase           b40c/graph/bfs/enactor_two_phase.cuh
//   EnactorHybrid             b40c/graph/bfs/enactor_hybrid.cuh
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "bfs_kernels.hpp"
#include "internal.hpp"

namespace b40c {

__global__ __launch_bounds__(256) void k_reset_seed(BfsArgs a, int ph, int32_t src, int mode) {
    reset_seed_body(a, ph, src, mode);
}

// Reset's label initialisation in one launch (MemsetKernel x 2 in the
// reference, csr_problem.cuh:723-745): depth and predecessor := -1, visited
// mask and the first two frontier bitmaps := 0.
__global__ __launch_bounds__(256) void k_reset(int32_t* depth, int32_t* pred, uint32_t* visited, uint32_t* f0,
                                               uint32_t* f1, int32_t n, int32_t words) {
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        depth[i] = -1;
        pred[i] = -1;
    }
    for (int32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
        visited[w] = 0;
        f0[w] = 0;
        f1[w] = 0;
    }
}

// Bit v = (vertex v has no neighbours), one thread per 32-vertex word.
__global__ __launch_bounds__(256) void k_isolated(const int32_t* offs, int32_t n, int32_t words, uint32_t* out) {
    const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint32_t bits = 0;
    const int32_t v0 = w * 32;
    for (int k = 0; k < 32 && v0 + k < n; ++k)
        if (offs[v0 + k + 1] == offs[v0 + k]) bits |= 1u << k;
    out[w] = bits;
}

// Last node of a captured level loop: advances the phase base and publishes
// the counters to the pinned host mirror, so the host's per-replay check
// needs no separate device-to-host copy.
__global__ void k_advance(BfsArgs a, int steps) {
    constexpr int kWords = (int)(sizeof(DevCounters) / 4);
    constexpr int kPhWord = (int)(offsetof(DevCounters, ph_base) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.ctr);
    uint32_t* dst = reinterpret_cast<uint32_t*>(a.hctr);
    for (int i = threadIdx.x; i < kWords; i += blockDim.x) {
        uint32_t v = src[i];
        if (i == kPhWord) {
            v += (uint32_t)steps;
            reinterpret_cast<uint32_t*>(a.ctr)[i] = v;
        }
        dst[i] = v;
    }
}

template <int KIND>
__global__ __launch_bounds__(BFS_THREADS) void k_phase(BfsArgs a, int ph, int32_t depth, int slot) {
    __shared__ LevelShared s;
    phase_body<KIND>(a, ph, depth, slot, s);
}

__global__ __launch_bounds__(BFS_THREADS) void k_do_level(BfsArgs a, int ph, int slot, int grid_tier) {
    __shared__ DoShared s;
    do_level_body(a, ph, slot, grid_tier, s);
}

template <int KIND>
__global__ __launch_bounds__(BFS_THREADS) void k_heavy(BfsArgs a, int ph, int32_t depth_next, int slot) {
    __shared__ typename HeavySharedOf<KIND>::type s;
    heavy_body<KIND>(a, ph, depth_next, slot, s);
}

template <int KIND>
__global__ __launch_bounds__(BFS_THREADS) void k_fused(BfsArgs a, int ph, int32_t depth, int max_levels,
                                                       uint32_t sat_limit) {
    __shared__ LevelShared s;
    __shared__ typename FusedHeavy<KIND>::shared hs;
    fused_body<KIND>(a, ph, depth, max_levels, sat_limit, s, hs);
}

__global__ __launch_bounds__(BFS_THREADS) void k_gbarrier_bench(GridBarrier bar, int iterations) {
    __shared__ uint32_t s_word;
    gbarrier_bench_body(bar, iterations, &s_word);
}

__global__ __launch_bounds__(BFS_THREADS) void k_gbarrier_check(GridBarrier bar, uint32_t* data, int words,
                                                                int iterations, unsigned long long* mismatches) {
    __shared__ uint32_t s_word;
    gbarrier_check_body(bar, data, words, iterations, mismatches, &s_word);
}

}  // namespace b40c

#include "bfs_host.inc"
