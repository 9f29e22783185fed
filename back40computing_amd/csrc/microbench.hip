// Per-stage BFS microbenchmarks for gfx950: a level-synchronous search split
// into its two stages, each launched twice per level, once as a bench pass
// (culls and reads only, on private copies of the visited mask and labels,
// counting survivors) and once for real.  Host-side timing with HIP events
// around every launch gives the per-level cost of each stage in isolation
// (test/bfs/microbench/enactor_gather_lookup.cuh:262-377 and
// enactor_cull.cuh:262-377 run the same alternation with per-kernel stats).
//
//   status lookup   vertex ids (edge frontier, with repeats) -> culled, newly
//                   labelled vertices as (row offset, row length) pairs
//                   (test/bfs/microbench/status_lookup/cta.cuh:450-527);
//   neighbour gather (row offset, row length) pairs -> every neighbour id,
//                   unculled (test/bfs/microbench/neighbor_gather/cta.cuh).
//
// The real lookup claims vertices exactly with the visited bitmask (one
// agent-scope atomicOr per candidate that passes the read-only bitmask test
// and a best-effort label mark) instead of the reference's compare-and-swap
// on the label, so the labels and the per-level vertex counts are
// deterministic.  The bench lookup of mode
// B40C_MB_CULL adds the label cull, a wave64 duplicate cull and the
// workgroup history cull (test/bfs/microbench/local_cull/cta.cuh:213-304,
// 525-533) and counts the survivors of each stage.
#include "bfs_device.hpp"
#include "internal.hpp"

#include <cstddef>

namespace b40c {
namespace {

constexpr int MB_THREADS = BFS_THREADS;
constexpr int MB_WAVE_HASH = 128;    // wave duplicate-cull table entries per wave
constexpr int LK_ITEMS = 4;          // lookup: ids per thread per pass
constexpr int LK_TILE = MB_THREADS * LK_ITEMS;

struct alignas(64) MbCounters {
    uint32_t pairs_len;              // lookup output: vertices with out-edges, as row ranges
    uint32_t discovered;             // lookup: vertices labelled this iteration
    uint32_t pass[3];                // bench lookup survivors: bitmask, label, local culls
    uint32_t overflow;
    uint32_t corrupt;
    uint32_t pad;
    unsigned long long next_edges;   // lookup: degree sum of the labelled vertices
    unsigned long long gather_sum;   // bench gather: sum of gathered neighbour ids
    unsigned long long scan_total;   // gather set-up: sum of the row lengths
};

struct LookupShared {
    HistoryHash hist;
    int32_t wave_hash[BFS_WAVES][MB_WAVE_HASH];
    uint32_t red[3][BFS_WAVES];
};

// Workgroup sums of up to three per-thread counts, one atomic per counter.
__device__ __forceinline__ void wg_add3(const uint32_t c[3], uint32_t* dst, uint32_t (*red)[BFS_WAVES], int nc) {
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    for (int k = 0; k < nc; ++k) {
        uint32_t x = c[k];
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
        if (lane == 0) red[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < nc) {
        uint32_t t = 0;
        for (int w = 0; w < BFS_WAVES; ++w) t += red[threadIdx.x][w];
        if (t) atomic_add_u32(&dst[threadIdx.x], t);
    }
}

// Workgroup sum of a 64-bit per-thread value into *dst (one atomic).
__device__ __forceinline__ void wg_add_u64(unsigned long long x, unsigned long long* dst,
                                           unsigned long long* red) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    if (lane_id() == 0) red[threadIdx.x / kWave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < BFS_WAVES; ++w) t += red[w];
        if (t) atomic_add_u64(dst, t);
    }
}

__global__ void k_mb_seed(const int32_t* __restrict__ offs, int32_t src, int32_t* labels, uint32_t* mask,
                          uint32_t* bench_mask, int32_t* pair_off, int32_t* pair_len, MbCounters* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        labels[src] = 0;
        mask[src >> 5] |= 1u << (src & 31);
        bench_mask[src >> 5] |= 1u << (src & 31);
        const int32_t deg = offs[src + 1] - offs[src];
        pair_off[0] = offs[src];
        pair_len[0] = deg;
        ctr->pairs_len = deg > 0 ? 1u : 0u;
        ctr->discovered = 1;
        ctr->next_edges = (unsigned long long)deg;
    }
}

// Real status lookup: bitmask test, best-effort label mark, exact claim; the
// row ranges of labelled vertices that have out-edges go to the pair queue
// (one global atomic per LDS flush), and the workgroup adds its discovered
// count and degree sum with one atomic each.  The label mark (plain load,
// plain store of the iteration) thins the duplicates before the atomic: a
// copy that finds the mark already set drops out, and every copy that set it
// goes on to the claim, so exactly one copy of each new vertex wins.
__global__ void __launch_bounds__(MB_THREADS) k_mb_lookup(const int32_t* __restrict__ ids, uint32_t n_in,
                                                          const int32_t* __restrict__ offs, int32_t n,
                                                          uint32_t* mask, int32_t* labels, int32_t iteration,
                                                          int32_t* pair_off, int32_t* pair_len, uint32_t pair_cap,
                                                          MbCounters* ctr) {
    __shared__ LdsPairQueue q;
    __shared__ uint32_t red32[3][BFS_WAVES];
    __shared__ unsigned long long red64[BFS_WAVES];
    if (threadIdx.x == 0) q.count = 0;
    __syncthreads();
    const GlobalQueue out{pair_off, pair_len, &ctr->pairs_len, pair_cap, &ctr->overflow};
    uint32_t found = 0;
    unsigned long long edges = 0;
    for (uint32_t base = blockIdx.x * LK_TILE; base < n_in; base += gridDim.x * LK_TILE) {
        // LK_ITEMS independent ids per thread, each stage over all of them,
        // so their memory accesses overlap.
        int32_t v[LK_ITEMS];
        bool keep[LK_ITEMS];
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) {
            const uint32_t i = base + r * MB_THREADS + threadIdx.x;
            v[r] = i < n_in ? ids[i] : -1;
            keep[r] = v[r] >= 0 && v[r] < n;
            if (i < n_in && !keep[r]) st_relaxed(&ctr->corrupt, 1u);
        }
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r)
            if (keep[r]) keep[r] = !visited_test(mask, v[r]);
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r)
            if (keep[r]) keep[r] = labels[v[r]] == -1;
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) {
            if (keep[r]) labels[v[r]] = iteration;
            if (keep[r]) keep[r] = claim_exact(mask, v[r]);
        }
        int32_t start[LK_ITEMS], len[LK_ITEMS];
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) {
            start[r] = keep[r] ? offs[v[r]] : 0;
            len[r] = keep[r] ? offs[v[r] + 1] - start[r] : 0;
            found += keep[r];
            edges += (unsigned long long)len[r];
        }
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) lq_append<true>(q, out, keep[r] && len[r] > 0, start[r], len[r]);
        lq_flush<true>(q, out, LQ_CAP / 2);
    }
    lq_flush<true>(q, out, 1);
    const uint32_t c[3] = {found, 0u, 0u};
    wg_add3(c, &ctr->discovered, red32, 1);
    wg_add_u64(edges, &ctr->next_edges, red64);
}

// Bench status lookup: culls only, on the bench mask (and, with the label
// cull, on a copy of the labels), counting the survivors of each stage.
//   cull = 0: bitmask cull with a best-effort (plain) bit store, then a read
//             of the survivor's label (status_lookup/cta.cuh:163-205);
//   cull = 1: bitmask cull, label cull (label read, plain store of the
//             iteration), wave duplicate cull, history cull
//             (local_cull/cta.cuh:525-533).
__global__ void __launch_bounds__(MB_THREADS) k_mb_lookup_bench(const int32_t* __restrict__ ids, uint32_t n_in,
                                                                int32_t n, uint32_t* bench_mask, int32_t* labels,
                                                                int32_t iteration, int32_t cull, MbCounters* ctr) {
    __shared__ LookupShared s;
    history_init(s.hist);
    __syncthreads();
    volatile int32_t* wh = s.wave_hash[threadIdx.x / kWave];
    const int lane = lane_id();
    uint32_t c[3] = {0u, 0u, 0u};
    for (uint32_t base = blockIdx.x * LK_TILE; base < n_in; base += gridDim.x * LK_TILE) {
        int32_t v[LK_ITEMS];
        bool keep[LK_ITEMS];
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) {
            const uint32_t i = base + r * MB_THREADS + threadIdx.x;
            v[r] = i < n_in ? ids[i] : -1;
            keep[r] = v[r] >= 0 && v[r] < n;
        }
        uint32_t word[LK_ITEMS];
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) word[r] = keep[r] ? bench_mask[v[r] >> 5] : 0u;
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) {
            if (keep[r]) {
                const uint32_t bit = 1u << (v[r] & 31);
                keep[r] = (word[r] & bit) == 0;
                if (keep[r]) bench_mask[v[r] >> 5] = word[r] | bit;
            }
            c[0] += keep[r];
        }
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r)
            if (keep[r]) keep[r] = labels[v[r]] == -1;
#pragma unroll
        for (int r = 0; r < LK_ITEMS; ++r) {
            if (keep[r] && cull) labels[v[r]] = iteration;
            c[1] += keep[r];
        }
        if (cull) {
#pragma unroll
            for (int r = 0; r < LK_ITEMS; ++r) {
                // Wave duplicate cull: lanes holding the same id race for one
                // slot, then for the right to keep it; exactly one copy of
                // each id that owns its slot survives.
                const int h = v[r] & (MB_WAVE_HASH - 1);
                if (keep[r]) wh[h] = v[r];
                if (keep[r] && wh[h] == v[r]) {
                    wh[h] = lane;
                    if (wh[h] != lane) keep[r] = false;
                }
                if (keep[r]) keep[r] = !history_seen(s.hist, v[r]);
                c[2] += keep[r];
            }
        }
    }
    __syncthreads();
    wg_add3(c, ctr->pass, s.red, cull ? 3 : 2);
}

// ---------------------------------------------------------------------------
// Load-balanced neighbour gather.  The exclusive scan of the row lengths
// gives every gathered neighbour its output position; the output range is
// cut into tiles of GT_TILE positions, k_mb_partition finds each tile's first
// row range by binary search, and a workgroup gathers a whole tile: it stages
// the tile's row ranges in LDS and maps each output position back to its row
// range by binary search there.  Every workgroup moves the same number of
// edges whatever the degree skew, columns are read and ids written with
// consecutive lanes on consecutive addresses, and nothing is reserved with
// atomics.  Row lengths must be positive (the lookup enqueues no empty rows),
// so a tile spans at most GT_TILE + 1 row ranges.
// ---------------------------------------------------------------------------
constexpr int GT_ITEMS = 8;
constexpr int GT_TILE = MB_THREADS * GT_ITEMS;

// part[t] = last row range i with scan[i] <= t * GT_TILE, t in [0, tiles);
// part[tiles] = n_pairs - 1.
__global__ void __launch_bounds__(MB_THREADS) k_mb_partition(const uint32_t* __restrict__ scan, uint32_t n_pairs,
                                                             uint32_t tiles, int32_t* part) {
    const uint32_t t = blockIdx.x * MB_THREADS + threadIdx.x;
    if (t > tiles) return;
    if (t == tiles) {
        part[t] = (int32_t)n_pairs - 1;
        return;
    }
    const uint64_t x = (uint64_t)t * GT_TILE;
    uint32_t lo = 0, hi = n_pairs;  // first index with scan > x lies in [lo, hi]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)scan[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    part[t] = (int32_t)lo - 1;
}

struct GatherTileShared {
    uint32_t scan[GT_TILE + 1];
    int32_t start[GT_TILE + 1];
    unsigned long long red[BFS_WAVES];
};

// bench = 1: the neighbour ids are only summed (one atomic per workgroup),
// nothing is written.
__global__ void __launch_bounds__(MB_THREADS) k_mb_gather(const int32_t* __restrict__ pair_off,
                                                          const uint32_t* __restrict__ scan,
                                                          const int32_t* __restrict__ part, uint32_t tiles,
                                                          uint64_t n_out, const int32_t* __restrict__ cols,
                                                          int32_t* out_ids, uint64_t out_cap, int32_t bench,
                                                          MbCounters* ctr) {
    __shared__ GatherTileShared s;
    const int tid = threadIdx.x;
    unsigned long long acc = 0;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t t0 = (uint64_t)t * GT_TILE;
        const uint64_t t1 = t0 + GT_TILE < n_out ? t0 + GT_TILE : n_out;
        const int32_t v0 = part[t];
        int32_t nv = part[t + 1] - v0 + 1;
        if (v0 < 0 || nv < 1 || nv > GT_TILE + 1) {
            // Not a partition of positive row lengths: flag it, touch nothing.
            if (tid == 0) st_relaxed(&ctr->corrupt, 1u);
            continue;
        }
        __syncthreads();
        for (int k = tid; k < nv; k += MB_THREADS) {
            s.scan[k] = scan[v0 + k];
            s.start[k] = pair_off[v0 + k];
        }
        __syncthreads();
        // Column index of each of this thread's positions (last staged range
        // with scan <= j; positions rise with r), then all loads, then all
        // stores, so GT_ITEMS loads are in flight per thread.
        int64_t col[GT_ITEMS];
        int lo = 0;
#pragma unroll
        for (int r = 0; r < GT_ITEMS; ++r) {
            const uint64_t j = t0 + (uint64_t)r * MB_THREADS + tid;
            col[r] = -1;
            if (j < t1) {
                int hi = nv;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if ((uint64_t)s.scan[mid] <= j) lo = mid;
                    else hi = mid;
                }
                col[r] = (int64_t)s.start[lo] + (int64_t)(j - s.scan[lo]);
            }
        }
        int32_t v[GT_ITEMS];
#pragma unroll
        for (int r = 0; r < GT_ITEMS; ++r) v[r] = col[r] >= 0 ? cols[col[r]] : 0;
#pragma unroll
        for (int r = 0; r < GT_ITEMS; ++r) {
            const uint64_t j = t0 + (uint64_t)r * MB_THREADS + tid;
            if (col[r] < 0) continue;
            if (bench) acc += (unsigned long long)(uint32_t)v[r];
            else if (j < out_cap) out_ids[j] = v[r];
        }
    }
    if (bench) {
        __syncthreads();
        wg_add_u64(acc, &ctr->gather_sum, s.red);
    } else if (blockIdx.x == 0 && tid == 0 && n_out > out_cap) {
        st_relaxed(&ctr->overflow, 1u);
    }
}

inline unsigned grid_for(uint64_t items, int cap) {
    uint64_t g = (items + LK_TILE - 1) / LK_TILE;
    if (g < 1) g = 1;
    if (g > (uint64_t)cap) g = (uint64_t)cap;
    return (unsigned)g;
}

struct Events {
    hipEvent_t e[8] = {};
    int made = 0;
    ~Events() {
        for (int i = 0; i < made; ++i) (void)hipEventDestroy(e[i]);
    }
    b40c_status create() {
        for (; made < 8; ++made) B40C_HIP(hipEventCreate(&e[made]));
        return B40C_OK;
    }
    float ms(int i, int j) const {
        float t = 0.f;
        if (hipEventElapsedTime(&t, e[i], e[j]) != hipSuccess) return -1.f;
        return t;
    }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
        if (s) (void)hipStreamDestroy(s);
    }
};

}  // namespace
}  // namespace b40c

using b40c::fail;

extern "C" b40c_status b40c_microbench_search(const b40c_graph* g, int32_t src, int32_t mode, int32_t* depth,
                                              b40c_mb_level* levels, int32_t max_levels, int32_t* num_levels) {
    using namespace b40c;
    if (!g || !num_levels || max_levels < 0 || (max_levels > 0 && !levels))
        return fail(B40C_ERR_INVALID_ARG, "b40c_microbench_search: bad arguments");
    if (mode != B40C_MB_GATHER_LOOKUP && mode != B40C_MB_CULL)
        return fail(B40C_ERR_INVALID_ARG, "b40c_microbench_search: unknown mode %d", mode);
    const int32_t n = g->nodes;
    if (src < 0 || src >= n)
        return fail(B40C_ERR_INVALID_ARG, "b40c_microbench_search: source %d outside [0, %d)", src, n);
    *num_levels = 0;
    DeviceInfo info;
    B40C_TRY(device_info(info));
    const int cap_grid = info.num_cus * 8;
    const size_t words = ((size_t)n + 31) / 32;
    const size_t m = (size_t)g->edges;
    const size_t max_tiles = (m + GT_TILE - 1) / GT_TILE;

    DeviceBuffer labels, bench_labels, mask, bench_mask, ids, pair_off, pair_len, pair_scan, part, partials,
        ctr_buf;
    B40C_TRY(labels.alloc((size_t)n * 4));
    B40C_TRY(mask.alloc(words * 4));
    B40C_TRY(bench_mask.alloc(words * 4));
    B40C_TRY(ids.alloc((m ? m : 1) * 4));
    B40C_TRY(pair_off.alloc((size_t)n * 4));
    B40C_TRY(pair_len.alloc((size_t)n * 4));
    B40C_TRY(pair_scan.alloc((size_t)n * 4));
    B40C_TRY(part.alloc((max_tiles + 1) * 4));
    B40C_TRY(partials.alloc((size_t)scan_u32_max_partials() * 8));
    B40C_TRY(ctr_buf.alloc(sizeof(MbCounters)));
    if (mode == B40C_MB_CULL) B40C_TRY(bench_labels.alloc((size_t)n * 4));
    StreamGuard sg;
    B40C_HIP(hipStreamCreate(&sg.s));
    hipStream_t st = sg.s;
    Events ev;
    B40C_TRY(ev.create());

    MbCounters* ctr = ctr_buf.as<MbCounters>();
    B40C_TRY(fill_u32(labels.as<uint32_t>(), 0xFFFFFFFFu, (size_t)n, st));
    B40C_HIP(hipMemsetAsync(mask.ptr, 0, words * 4, st));
    B40C_HIP(hipMemsetAsync(bench_mask.ptr, 0, words * 4, st));
    B40C_HIP(hipMemsetAsync(ctr, 0, sizeof(MbCounters), st));
    hipLaunchKernelGGL(k_mb_seed, dim3(1), dim3(64), 0, st, g->d_offsets(), src, labels.as<int32_t>(),
                       mask.as<uint32_t>(), bench_mask.as<uint32_t>(), pair_off.as<int32_t>(),
                       pair_len.as<int32_t>(), ctr);
    B40C_HIP(hipGetLastError());

    MbCounters h{};
    auto read_ctr = [&]() -> b40c_status {
        B40C_HIP(hipMemcpyAsync(&h, ctr, sizeof(MbCounters), hipMemcpyDeviceToHost, st));
        B40C_HIP(hipStreamSynchronize(st));
        if (h.overflow) return fail(B40C_ERR_QUEUE_OVERFLOW, "b40c_microbench_search: frontier queue overflow");
        if (h.corrupt) return fail(B40C_ERR_HIP, "b40c_microbench_search: frontier or partition out of range");
        return B40C_OK;
    };
    B40C_TRY(read_ctr());

    b40c_mb_level rec{};
    rec.iteration = 0;
    rec.bitmask_pass = rec.label_pass = rec.local_pass = -1;
    rec.discovered = 1;
    int32_t count = 0;
    for (int32_t it = 0;; ++it) {
        // Neighbour gather of the vertices labelled `it`: scan of the row
        // lengths and tile partition, then the bench pass and the real pass.
        const uint32_t n_pairs = h.pairs_len;
        const uint64_t n_out = h.next_edges;
        rec.gathered = (int64_t)n_out;
        if (n_pairs > 0) {
            if (n_out > m) return fail(B40C_ERR_HIP, "b40c_microbench_search: degree sum exceeds the edge count");
            const uint32_t tiles = (uint32_t)((n_out + GT_TILE - 1) / GT_TILE);
            const unsigned gg = (unsigned)(tiles < (uint32_t)cap_grid ? tiles : (uint32_t)cap_grid);
            B40C_HIP(hipMemsetAsync(&ctr->gather_sum, 0, 16, st));
            B40C_HIP(hipEventRecord(ev.e[0], st));
            B40C_TRY(scan_u32_async(pair_len.as<uint32_t>(), pair_scan.as<uint32_t>(), n_pairs, true,
                                    partials.as<uint64_t>(), (uint64_t*)&ctr->scan_total, st));
            hipLaunchKernelGGL(k_mb_partition, dim3((tiles + MB_THREADS) / MB_THREADS), dim3(MB_THREADS), 0, st,
                               pair_scan.as<uint32_t>(), n_pairs, tiles, part.as<int32_t>());
            B40C_HIP(hipEventRecord(ev.e[1], st));
            hipLaunchKernelGGL(k_mb_gather, dim3(gg), dim3(MB_THREADS), 0, st, pair_off.as<int32_t>(),
                               pair_scan.as<uint32_t>(), part.as<int32_t>(), tiles, n_out, g->d_cols(),
                               ids.as<int32_t>(), (uint64_t)m, 1, ctr);
            B40C_HIP(hipEventRecord(ev.e[2], st));
            hipLaunchKernelGGL(k_mb_gather, dim3(gg), dim3(MB_THREADS), 0, st, pair_off.as<int32_t>(),
                               pair_scan.as<uint32_t>(), part.as<int32_t>(), tiles, n_out, g->d_cols(),
                               ids.as<int32_t>(), (uint64_t)m, 0, ctr);
            B40C_HIP(hipEventRecord(ev.e[3], st));
            B40C_HIP(hipGetLastError());
            B40C_TRY(read_ctr());
            if (h.scan_total != n_out)
                return fail(B40C_ERR_HIP, "b40c_microbench_search: row lengths sum to %llu, expected %llu",
                            (unsigned long long)h.scan_total, (unsigned long long)n_out);
            rec.gather_checksum = h.gather_sum;
            rec.offsets_ms = ev.ms(0, 1);
            rec.bench_gather_ms = ev.ms(1, 2);
            rec.gather_ms = ev.ms(2, 3);
        }
        if (count < max_levels) levels[count] = rec;
        ++count;
        if (n_out == 0) break;

        // Status lookup labelling `it + 1`: bench, then real.
        const uint32_t n_ids = (uint32_t)n_out;
        rec = b40c_mb_level{};
        rec.iteration = it + 1;
        rec.lookup_in = n_ids;
        B40C_HIP(hipMemsetAsync(ctr, 0, offsetof(MbCounters, overflow), st));
        B40C_HIP(hipMemsetAsync(&ctr->next_edges, 0, 8, st));
        if (mode == B40C_MB_CULL)
            B40C_HIP(hipMemcpyAsync(bench_labels.ptr, labels.ptr, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        const unsigned gl = grid_for(n_ids, cap_grid);
        int32_t* blab = mode == B40C_MB_CULL ? bench_labels.as<int32_t>() : labels.as<int32_t>();
        B40C_HIP(hipEventRecord(ev.e[4], st));
        hipLaunchKernelGGL(k_mb_lookup_bench, dim3(gl), dim3(MB_THREADS), 0, st, ids.as<int32_t>(), n_ids, n,
                           bench_mask.as<uint32_t>(), blab, it + 1, mode == B40C_MB_CULL ? 1 : 0, ctr);
        B40C_HIP(hipEventRecord(ev.e[5], st));
        hipLaunchKernelGGL(k_mb_lookup, dim3(gl), dim3(MB_THREADS), 0, st, ids.as<int32_t>(), n_ids,
                           g->d_offsets(), n, mask.as<uint32_t>(), labels.as<int32_t>(), it + 1,
                           pair_off.as<int32_t>(), pair_len.as<int32_t>(), (uint32_t)n, ctr);
        B40C_HIP(hipEventRecord(ev.e[6], st));
        B40C_HIP(hipGetLastError());
        B40C_TRY(read_ctr());
        rec.bitmask_pass = h.pass[0];
        rec.label_pass = h.pass[1];
        rec.local_pass = mode == B40C_MB_CULL ? (int64_t)h.pass[2] : -1;
        rec.discovered = h.discovered;
        rec.bench_lookup_ms = ev.ms(4, 5);
        rec.lookup_ms = ev.ms(5, 6);
        if (h.discovered == 0) {
            if (count < max_levels) levels[count] = rec;
            ++count;
            break;
        }
    }
    if (depth) {
        B40C_HIP(hipMemcpyAsync(depth, labels.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        B40C_HIP(hipStreamSynchronize(st));
    }
    *num_levels = count;
    return B40C_OK;
}
