// Device-resident CSR graphs and their GPU builders: seeded RMAT and 3-D
// lattice generators, packed-key radix-sort CSR construction (sort ->
// unique -> row offsets), and COO -> CSR with the reference's stable
// row-only ordering.
//
// Reference counterparts: CsrGraph + FromCoo (b40c/graph/csr_graph.cuh:50-152),
// BuildRmatGraph (graph/builder/rmat.cuh:127-206), BuildGrid3dGraph
// (graph/builder/grid3d.cuh:53-127).  The reference builds on the host with
// rand() and std::stable_sort; here every step runs on the GPU with a
// counter-based generator so the stream is identical for any launch shape.
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace b40c {

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint32_t prob_threshold(double p) {
    double t = p * 4294967296.0;
    if (!(t > 0.0)) return 0u;
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)t;
}

void graph_retain(b40c_graph* g) {
    if (g) g->refs.fetch_add(1);
}

void graph_release(b40c_graph* g) {
    if (g && g->refs.fetch_sub(1) == 1) delete g;
}

namespace {
__global__ void k_check_symmetric(const int32_t* offs, const int32_t* cols, int32_t n, uint32_t* flags);
__global__ void k_max_degree(const int32_t* offs, int32_t n, uint32_t* out);
}

b40c_status graph_max_degree(b40c_graph* g, int32_t& max_degree) {
    int known = g->max_degree.load();
    if (known < 0) {
        DeviceBuffer out;
        B40C_TRY(out.alloc(sizeof(uint32_t)));
        B40C_HIP(hipMemset(out.ptr, 0, sizeof(uint32_t)));
        hipLaunchKernelGGL(k_max_degree, dim3(1024), dim3(256), 0, nullptr, g->d_offsets(), g->nodes,
                           out.as<uint32_t>());
        B40C_HIP(hipGetLastError());
        uint32_t h = 0;
        B40C_HIP(hipMemcpy(&h, out.ptr, sizeof(h), hipMemcpyDeviceToHost));
        known = (int)h;
        g->max_degree.store(known);
    }
    max_degree = known;
    return B40C_OK;
}

b40c_status graph_symmetric(b40c_graph* g, bool& symmetric) {
    int known = g->symmetric.load();
    if (known < 0) {
        DeviceBuffer flags;
        B40C_TRY(flags.alloc(2 * sizeof(uint32_t)));
        B40C_HIP(hipMemset(flags.ptr, 0, 2 * sizeof(uint32_t)));
        if (g->edges > 0) {
            hipLaunchKernelGGL(k_check_symmetric, dim3(4096), dim3(256), 0, nullptr, g->d_offsets(), g->d_cols(),
                               g->nodes, flags.as<uint32_t>());
            B40C_HIP(hipGetLastError());
        }
        uint32_t h[2] = {0, 0};
        B40C_HIP(hipMemcpy(h, flags.ptr, sizeof(h), hipMemcpyDeviceToHost));
        known = (h[0] == 0 && h[1] == 0) ? 1 : 0;
        g->symmetric.store(known);
    }
    symmetric = known == 1;
    return B40C_OK;
}

namespace {

constexpr int GB_THREADS = 256;

// One sampled edge per thread (grid-stride).  Draw l of edge i is 32-bit half
// (l & 1) of splitmix64(smix + 16*i + l/2); quadrant a keeps (u,v), b adds the
// step to the column, c to the row, d to both (ChoosePartition,
// graph/builder/rmat.cuh:55-81).  Both directions are emitted as packed
// (src << scale) | dst keys; self loops become the sentinel key.
__global__ __launch_bounds__(GB_THREADS) void k_rmat_keys(int scale, int64_t m, uint32_t tA, uint32_t tAB,
                                                          uint32_t tABC, uint64_t smix, uint64_t* keys,
                                                          uint64_t sentinel) {
    for (int64_t i = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; i < m; i += (int64_t)gridDim.x * GB_THREADS) {
        uint32_t u = 0, v = 0;
        uint64_t x = 0;
        for (int l = 0; l < scale; ++l) {
            if ((l & 1) == 0) x = splitmix64(smix + 16ull * (uint64_t)i + (uint64_t)(l >> 1));
            const uint32_t r = (l & 1) ? (uint32_t)(x >> 32) : (uint32_t)x;
            const uint32_t step = 1u << (scale - 1 - l);
            if (r < tA) {
            } else if (r < tAB) {
                v += step;
            } else if (r < tABC) {
                u += step;
            } else {
                u += step;
                v += step;
            }
        }
        if (u == v) {
            keys[2 * i] = sentinel;
            keys[2 * i + 1] = sentinel;
        } else {
            keys[2 * i] = ((uint64_t)u << scale) | v;
            keys[2 * i + 1] = ((uint64_t)v << scale) | u;
        }
    }
}

// Relabelling keys of the lattice: random high bits, original id in the low bits.
__global__ __launch_bounds__(GB_THREADS) void k_perm_keys(int64_t n, uint64_t lowmask, uint64_t smix, uint64_t* keys) {
    for (int64_t o = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; o < n; o += (int64_t)gridDim.x * GB_THREADS)
        keys[o] = (splitmix64(smix + (uint64_t)o) & ~lowmask) | (uint64_t)o;
}

__global__ __launch_bounds__(GB_THREADS) void k_perm_from_sorted(const uint64_t* sorted, int64_t n, uint64_t lowmask,
                                                                 int32_t* perm) {
    for (int64_t r = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; r < n; r += (int64_t)gridDim.x * GB_THREADS)
        perm[sorted[r] & lowmask] = (int32_t)r;
}

// Six neighbour keys per original vertex in the reference's neighbour order
// (k-1, k+1, j-1, j+1, i-1, i+1; grid3d.cuh:79-113), no self loop.
__global__ __launch_bounds__(GB_THREADS) void k_lattice_keys(int w, const int32_t* perm, int shift, uint64_t* keys,
                                                             uint64_t sentinel) {
    const int64_t n = (int64_t)w * w * w;
    const int64_t ww = (int64_t)w * w;
    for (int64_t o = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; o < n; o += (int64_t)gridDim.x * GB_THREADS) {
        const int64_t i = o / ww, j = (o / w) % w, k = o % w;
        const uint64_t p = (uint64_t)perm[o] << shift;
        uint64_t* out = keys + 6 * o;
        out[0] = (k - 1 >= 0) ? (p | (uint64_t)perm[o - 1]) : sentinel;
        out[1] = (k + 1 < w) ? (p | (uint64_t)perm[o + 1]) : sentinel;
        out[2] = (j - 1 >= 0) ? (p | (uint64_t)perm[o - w]) : sentinel;
        out[3] = (j + 1 < w) ? (p | (uint64_t)perm[o + w]) : sentinel;
        out[4] = (i - 1 >= 0) ? (p | (uint64_t)perm[o - ww]) : sentinel;
        out[5] = (i + 1 < w) ? (p | (uint64_t)perm[o + ww]) : sentinel;
    }
}

__device__ __forceinline__ bool keep_key(const uint64_t* keys, int64_t i, uint64_t sentinel) {
    const uint64_t k = keys[i];
    return k != sentinel && (i == 0 || keys[i - 1] != k);
}

// Pass 1 of the unique compaction: kept keys per workgroup chunk.
__global__ __launch_bounds__(GB_THREADS) void k_unique_count(const uint64_t* keys, int64_t nkeys, uint64_t sentinel,
                                                             int64_t chunk, uint32_t* wg_counts) {
    __shared__ int s_wave[GB_THREADS / kWave];
    const int64_t begin = (int64_t)blockIdx.x * chunk;
    const int64_t end = min(nkeys, begin + chunk);
    int c = 0;
    for (int64_t i = begin + threadIdx.x; i < end; i += GB_THREADS) c += keep_key(keys, i, sentinel) ? 1 : 0;
    int total;
    (void)wg_exclusive_scan<GB_THREADS>(c, total, s_wave);
    if (threadIdx.x == 0) wg_counts[blockIdx.x] = (uint32_t)total;
}

// Exclusive scan of up to a few thousand workgroup counts (one workgroup).
__global__ __launch_bounds__(GB_THREADS) void k_scan_counts(uint32_t* counts, int G, uint32_t* total_out) {
    __shared__ int s_wave[GB_THREADS / kWave];
    uint32_t carry = 0;
    for (int base = 0; base < G; base += GB_THREADS) {
        const int i = base + threadIdx.x;
        const int v = (i < G) ? (int)counts[i] : 0;
        int tot;
        const int ex = wg_exclusive_scan<GB_THREADS>(v, tot, s_wave);
        if (i < G) counts[i] = carry + (uint32_t)ex;
        carry += (uint32_t)tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

// Pass 2: write the kept keys' columns and rows in order.
__global__ __launch_bounds__(GB_THREADS) void k_unique_scatter(const uint64_t* keys, int64_t nkeys, uint64_t sentinel,
                                                               int64_t chunk, const uint32_t* wg_offsets, int shift,
                                                               int32_t* cols, uint32_t* rows) {
    __shared__ int s_wave[GB_THREADS / kWave];
    const int64_t begin = (int64_t)blockIdx.x * chunk;
    const int64_t end = min(nkeys, begin + chunk);
    const uint64_t mask = (1ull << shift) - 1;
    uint32_t carry = wg_offsets[blockIdx.x];
    for (int64_t base = begin; base < end; base += GB_THREADS) {
        const int64_t i = base + threadIdx.x;
        const bool keep = (i < end) && keep_key(keys, i, sentinel);
        int tot;
        const int ex = wg_exclusive_scan<GB_THREADS>(keep ? 1 : 0, tot, s_wave);
        if (keep) {
            const uint64_t k = keys[i];
            cols[carry + ex] = (int32_t)(k & mask);
            rows[carry + ex] = (uint32_t)(k >> shift);
        }
        carry += (uint32_t)tot;
    }
}

// Row offsets from the row of every edge (rows ascending): edge p opens rows
// (rows[p-1], rows[p]]; the last edge closes the trailing rows.  Same fill
// rule as CsrGraph::FromCoo (csr_graph.cuh:127-147).
__global__ __launch_bounds__(GB_THREADS) void k_row_offsets(const uint32_t* rows, int64_t E, int32_t n,
                                                            int32_t* offsets) {
    for (int64_t p = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; p < E; p += (int64_t)gridDim.x * GB_THREADS) {
        const int64_t r = rows[p];
        const int64_t rp = (p == 0) ? -1 : (int64_t)rows[p - 1];
        for (int64_t row = rp + 1; row <= r; ++row) offsets[row] = (int32_t)p;
        if (p == E - 1)
            for (int64_t row = r + 1; row <= n; ++row) offsets[row] = (int32_t)E;
    }
}

// Symmetry check used before bottom-up steps: one wave per vertex u; lanes
// stride u's columns, check that they are strictly increasing and that each
// neighbour v lists u (binary search in v's sorted columns).  flags[0] counts
// unsorted rows or out-of-range ids, flags[1] missing reverse edges.
__global__ __launch_bounds__(GB_THREADS) void k_check_symmetric(const int32_t* offs, const int32_t* cols, int32_t n,
                                                                uint32_t* flags) {
    const int lane = threadIdx.x % kWave;
    const int64_t wave0 = ((int64_t)blockIdx.x * GB_THREADS + threadIdx.x) / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * GB_THREADS / kWave;
    uint32_t bad_order = 0, missing = 0;
    for (int64_t u = wave0; u < n; u += nwaves) {
        const int32_t s = offs[u], e = offs[u + 1];
        for (int32_t k = s + lane; k < e; k += kWave) {
            const int32_t v = cols[k];
            if (v < 0 || v >= n) {
                ++bad_order;
                continue;
            }
            if (k + 1 < e && cols[k + 1] <= v) ++bad_order;
            int32_t lo = offs[v], hi = offs[v + 1];
            while (lo < hi) {
                const int32_t mid = lo + (hi - lo) / 2;
                if (cols[mid] < (int32_t)u)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo >= offs[v + 1] || cols[lo] != (int32_t)u) ++missing;
        }
    }
    if (bad_order) atomic_add_u32(&flags[0], bad_order);
    if (missing) atomic_add_u32(&flags[1], missing);
}

// Largest out-degree: grid-stride over vertices, wave maximum, one atomic
// per wave.
__global__ __launch_bounds__(GB_THREADS) void k_max_degree(const int32_t* offs, int32_t n, uint32_t* out) {
    uint32_t m = 0;
    for (int64_t v = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; v < n; v += (int64_t)gridDim.x * GB_THREADS)
        m = max(m, (uint32_t)(offs[v + 1] - offs[v]));
    for (int d = kWave / 2; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, kWave));
    if (threadIdx.x % kWave == 0 && m) atomicMax(out, m);
}

__global__ void k_u32_to_i32_copy(const uint32_t* src, int32_t* dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (int32_t)src[i];
}

unsigned grid_for(int64_t n, int64_t cap = 16384) {
    int64_t b = (n + GB_THREADS - 1) / GB_THREADS;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// Sorted packed keys -> deduplicated CSR in g (sentinel keys dropped).
b40c_status csr_from_sorted_keys(const uint64_t* keys, int64_t nkeys, int shift, uint64_t sentinel, int32_t n,
                                 b40c_graph* g, hipStream_t stream) {
    DeviceInfo info;
    B40C_TRY(device_info(info));
    int64_t G = (int64_t)info.num_cus * 4;
    int64_t chunk = (nkeys + G - 1) / G;
    if (chunk < 1) chunk = 1;
    G = (nkeys + chunk - 1) / chunk;
    if (G < 1) G = 1;
    DeviceBuffer counts, total;
    B40C_TRY(counts.alloc(G * sizeof(uint32_t)));
    B40C_TRY(total.alloc(sizeof(uint32_t)));
    hipLaunchKernelGGL(k_unique_count, dim3((unsigned)G), dim3(GB_THREADS), 0, stream, keys, nkeys, sentinel, chunk,
                       counts.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(GB_THREADS), 0, stream, counts.as<uint32_t>(), (int)G,
                       total.as<uint32_t>());
    B40C_HIP(hipGetLastError());
    uint32_t E = 0;
    B40C_HIP(hipMemcpyAsync(&E, total.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    B40C_HIP(hipStreamSynchronize(stream));
    if ((int64_t)E > 0x7FFFFFFFll) return fail(B40C_ERR_UNSUPPORTED, "CSR with %u edges exceeds int32 offsets", E);
    g->nodes = n;
    g->edges = E;
    DeviceBuffer rows;
    B40C_TRY(g->cols.alloc((size_t)E * sizeof(int32_t)));
    B40C_TRY(rows.alloc((size_t)E * sizeof(uint32_t)));
    B40C_TRY(g->offsets.alloc(((size_t)n + 1) * sizeof(int32_t)));
    hipLaunchKernelGGL(k_unique_scatter, dim3((unsigned)G), dim3(GB_THREADS), 0, stream, keys, nkeys, sentinel, chunk,
                       counts.as<uint32_t>(), shift, g->cols.as<int32_t>(), rows.as<uint32_t>());
    if (E == 0) {
        B40C_HIP(hipMemsetAsync(g->offsets.ptr, 0, ((size_t)n + 1) * sizeof(int32_t), stream));
    } else {
        hipLaunchKernelGGL(k_row_offsets, dim3(grid_for(E)), dim3(GB_THREADS), 0, stream, rows.as<uint32_t>(),
                           (int64_t)E, n, g->offsets.as<int32_t>());
    }
    B40C_HIP(hipGetLastError());
    B40C_HIP(hipStreamSynchronize(stream));
    return B40C_OK;
}

// COO edge e -> packed key (row << shift) | col, plus its reverse at m + e
// when symmetrizing; dropped self-loops become the sentinel.
__global__ __launch_bounds__(GB_THREADS) void k_pack_coo(const int32_t* rows, const int32_t* cols, int64_t m,
                                                         int shift, int drop_loops, int symmetrize, uint64_t sentinel,
                                                         uint64_t* keys) {
    for (int64_t e = (int64_t)blockIdx.x * GB_THREADS + threadIdx.x; e < m; e += (int64_t)gridDim.x * GB_THREADS) {
        const uint64_t r = (uint32_t)rows[e], c = (uint32_t)cols[e];
        const bool loop = drop_loops && r == c;
        keys[e] = loop ? sentinel : ((r << shift) | c);
        if (symmetrize) keys[m + e] = loop ? sentinel : ((c << shift) | r);
    }
}

// COO -> CSR with sorted, unique columns per row: pack, GPU radix sort of the
// packed keys, consecutive removal, row offsets (the steps around the sort
// that the reference does with consecutive_removal::Enactor::Trim and
// scan::Enactor::Scan).
b40c_status build_from_coo_unique(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, int32_t flags,
                                  b40c_graph* g) {
    if (n < 1 || m < 0 || (m > 0 && (!rows || !cols)) || (flags & ~(B40C_COO_DROP_SELF_LOOPS | B40C_COO_SYMMETRIZE)))
        return fail(B40C_ERR_INVALID_ARG, "from_coo_unique: bad arguments");
    for (int64_t e = 0; e < m; ++e)
        if (rows[e] < 0 || rows[e] >= n || cols[e] < 0 || cols[e] >= n)
            return fail(B40C_ERR_INVALID_ARG, "from_coo_unique: edge %lld (%d,%d) out of range [0,%d)",
                        (long long)e, rows[e], cols[e], n);
    const bool sym = (flags & B40C_COO_SYMMETRIZE) != 0;
    const int64_t nkeys = sym ? 2 * m : m;
    if (nkeys > 0xFFFFFFFFll) return fail(B40C_ERR_UNSUPPORTED, "from_coo_unique: %lld keys exceed the sorter",
                                          (long long)nkeys);
    hipStream_t stream = nullptr;
    const int shift = std::max(1, ceil_log2_u64((uint64_t)n));
    const uint64_t sentinel = 1ull << (2 * shift);  // above every real key
    DeviceBuffer r_d, c_d, k0, k1;
    B40C_TRY(r_d.alloc((size_t)std::max<int64_t>(m, 1) * 4));
    B40C_TRY(c_d.alloc((size_t)std::max<int64_t>(m, 1) * 4));
    B40C_TRY(k0.alloc((size_t)std::max<int64_t>(nkeys, 1) * 8));
    B40C_TRY(k1.alloc((size_t)std::max<int64_t>(nkeys, 1) * 8));
    if (m > 0) {
        B40C_HIP(hipMemcpy(r_d.ptr, rows, (size_t)m * 4, hipMemcpyHostToDevice));
        B40C_HIP(hipMemcpy(c_d.ptr, cols, (size_t)m * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_pack_coo, dim3(grid_for(m)), dim3(GB_THREADS), 0, stream, r_d.as<int32_t>(),
                           c_d.as<int32_t>(), m, shift, (flags & B40C_COO_DROP_SELF_LOOPS) ? 1 : 0, sym ? 1 : 0,
                           sentinel, k0.as<uint64_t>());
        B40C_HIP(hipGetLastError());
    }
    int32_t sel = 0;
    if (nkeys > 0) B40C_TRY(radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), nkeys, 0, 2 * shift + 1, &sel, stream));
    const uint64_t* sorted = sel ? k1.as<uint64_t>() : k0.as<uint64_t>();
    B40C_TRY(csr_from_sorted_keys(sorted, nkeys, shift, sentinel, n, g, stream));
    if (sym) g->symmetric.store(1);
    return B40C_OK;
}

b40c_status build_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, b40c_graph* g) {
    if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(B40C_ERR_INVALID_ARG, "rmat: bad parameters (scale=%d ef=%d a=%g b=%g c=%g)", scale, edge_factor,
                    a, b, c);
    const int32_t n = 1 << scale;
    const int64_t m = (int64_t)edge_factor << scale;
    const int64_t nkeys = 2 * m;
    if (nkeys > 0xFFFFFFFFll) return fail(B40C_ERR_UNSUPPORTED, "rmat: %lld keys exceed the sorter", (long long)nkeys);
    hipStream_t stream = nullptr;
    DeviceBuffer k0, k1;
    B40C_TRY(k0.alloc((size_t)nkeys * sizeof(uint64_t)));
    B40C_TRY(k1.alloc((size_t)nkeys * sizeof(uint64_t)));
    const uint64_t sentinel = (scale * 2 >= 64) ? ~0ull : ((1ull << (2 * scale)) - 1);
    hipLaunchKernelGGL(k_rmat_keys, dim3(grid_for(m)), dim3(GB_THREADS), 0, stream, scale, m, prob_threshold(a),
                       prob_threshold(a + b), prob_threshold(a + b + c), splitmix64(seed), k0.as<uint64_t>(),
                       sentinel);
    B40C_HIP(hipGetLastError());
    int32_t sel = 0;
    B40C_TRY(radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), nkeys, 0, 2 * scale, &sel, stream));
    const uint64_t* sorted = sel ? k1.as<uint64_t>() : k0.as<uint64_t>();
    B40C_TRY(csr_from_sorted_keys(sorted, nkeys, scale, sentinel, n, g, stream));
    g->symmetric.store(1);  // both directions of every sampled edge were emitted
    return B40C_OK;
}

b40c_status build_lattice(int w, uint64_t seed, int32_t* h_perm, b40c_graph* g) {
    if (w < 1 || (int64_t)w * w * w > (1ll << 28))
        return fail(B40C_ERR_INVALID_ARG, "lattice: width %d out of range", w);
    const int64_t n = (int64_t)w * w * w;
    const int bits = ceil_log2_u64((uint64_t)n);
    const uint64_t lowmask = (1ull << bits) - 1;
    hipStream_t stream = nullptr;
    const int64_t nkeys = 6 * n;
    DeviceBuffer k0, k1, perm;
    B40C_TRY(k0.alloc((size_t)nkeys * sizeof(uint64_t)));
    B40C_TRY(k1.alloc((size_t)nkeys * sizeof(uint64_t)));
    B40C_TRY(perm.alloc((size_t)n * sizeof(int32_t)));
    hipLaunchKernelGGL(k_perm_keys, dim3(grid_for(n)), dim3(GB_THREADS), 0, stream, n, lowmask, splitmix64(seed),
                       k0.as<uint64_t>());
    int32_t sel = 0;
    B40C_TRY(radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), n, 0, 64, &sel, stream));
    hipLaunchKernelGGL(k_perm_from_sorted, dim3(grid_for(n)), dim3(GB_THREADS), 0, stream,
                       sel ? k1.as<uint64_t>() : k0.as<uint64_t>(), n, lowmask, perm.as<int32_t>());
    const uint64_t sentinel = (1ull << (2 * bits)) - 1;
    hipLaunchKernelGGL(k_lattice_keys, dim3(grid_for(n)), dim3(GB_THREADS), 0, stream, w, perm.as<int32_t>(), bits,
                       k0.as<uint64_t>(), sentinel);
    B40C_HIP(hipGetLastError());
    B40C_TRY(radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), nkeys, 0, 2 * bits, &sel, stream));
    B40C_TRY(csr_from_sorted_keys(sel ? k1.as<uint64_t>() : k0.as<uint64_t>(), nkeys, bits, sentinel, (int32_t)n, g,
                                  stream));
    if (h_perm) B40C_HIP(hipMemcpy(h_perm, perm.ptr, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
    g->symmetric.store(1);  // every lattice edge is emitted from both endpoints
    return B40C_OK;
}

b40c_status build_from_coo(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, b40c_graph* g) {
    if (n < 1 || m < 0 || m > 0x7FFFFFFFll || (m > 0 && (!rows || !cols)))
        return fail(B40C_ERR_INVALID_ARG, "from_coo: bad arguments");
    for (int64_t e = 0; e < m; ++e)
        if (rows[e] < 0 || rows[e] >= n || cols[e] < 0 || cols[e] >= n)
            return fail(B40C_ERR_INVALID_ARG, "from_coo: edge %lld (%d,%d) out of range [0,%d)", (long long)e,
                        rows[e], cols[e], n);
    hipStream_t stream = nullptr;
    g->nodes = n;
    g->edges = m;
    B40C_TRY(g->offsets.alloc(((size_t)n + 1) * sizeof(int32_t)));
    B40C_TRY(g->cols.alloc((size_t)m * sizeof(int32_t)));
    if (m == 0) {
        B40C_HIP(hipMemset(g->offsets.ptr, 0, ((size_t)n + 1) * sizeof(int32_t)));
        return B40C_OK;
    }
    DeviceBuffer k0, k1, v0, v1;
    B40C_TRY(k0.alloc((size_t)m * 4));
    B40C_TRY(k1.alloc((size_t)m * 4));
    B40C_TRY(v0.alloc((size_t)m * 4));
    B40C_TRY(v1.alloc((size_t)m * 4));
    B40C_HIP(hipMemcpy(k0.ptr, rows, (size_t)m * 4, hipMemcpyHostToDevice));
    B40C_HIP(hipMemcpy(v0.ptr, cols, (size_t)m * 4, hipMemcpyHostToDevice));
    int32_t sel = 0;
    const int bits = ceil_log2_u64((uint64_t)n);
    B40C_TRY(radix_sort_pairs_u32(k0.as<uint32_t>(), k1.as<uint32_t>(), v0.as<uint32_t>(), v1.as<uint32_t>(), m, 0,
                                  bits, &sel, stream));
    const uint32_t* srows = sel ? k1.as<uint32_t>() : k0.as<uint32_t>();
    const uint32_t* scols = sel ? v1.as<uint32_t>() : v0.as<uint32_t>();
    hipLaunchKernelGGL(k_u32_to_i32_copy, dim3(grid_for(m)), dim3(GB_THREADS), 0, stream, scols,
                       g->cols.as<int32_t>(), m);
    hipLaunchKernelGGL(k_row_offsets, dim3(grid_for(m)), dim3(GB_THREADS), 0, stream, srows, m, n,
                       g->offsets.as<int32_t>());
    B40C_HIP(hipGetLastError());
    B40C_HIP(hipStreamSynchronize(stream));
    return B40C_OK;
}

}  // namespace
}  // namespace b40c

using b40c::fail;

extern "C" {

b40c_status b40c_graph_from_host_csr(int32_t nodes, int64_t edges, const int32_t* row_offsets,
                                     const int32_t* column_indices, b40c_graph** out) {
    if (!out || nodes < 1 || edges < 0 || edges > 0x7FFFFFFFll || !row_offsets || (edges > 0 && !column_indices))
        return fail(B40C_ERR_INVALID_ARG, "b40c_graph_from_host_csr: bad arguments");
    if (row_offsets[0] != 0 || row_offsets[nodes] != edges)
        return fail(B40C_ERR_INVALID_ARG, "b40c_graph_from_host_csr: offsets[0]=%d offsets[n]=%d edges=%lld",
                    row_offsets[0], row_offsets[nodes], (long long)edges);
    // The kernels use neighbour ids and row ranges as addresses, so a CSR
    // whose offsets decrease or whose column ids leave [0, n) is refused
    // here instead of turning into out-of-bounds device accesses.
    for (int32_t v = 0; v < nodes; ++v)
        if (row_offsets[v + 1] < row_offsets[v])
            return fail(B40C_ERR_INVALID_ARG, "b40c_graph_from_host_csr: offsets decrease at vertex %d (%d > %d)", v,
                        row_offsets[v], row_offsets[v + 1]);
    for (int64_t e = 0; e < edges; ++e)
        if ((uint32_t)column_indices[e] >= (uint32_t)nodes)
            return fail(B40C_ERR_INVALID_ARG, "b40c_graph_from_host_csr: column id %d at entry %lld is out of range [0,%d)",
                        column_indices[e], (long long)e, nodes);
    b40c_graph* g = new b40c_graph();
    g->nodes = nodes;
    g->edges = edges;
    b40c_status s = g->offsets.alloc(((size_t)nodes + 1) * sizeof(int32_t));
    if (s == B40C_OK) s = g->cols.alloc((size_t)edges * sizeof(int32_t));
    if (s == B40C_OK) {
        hipError_t e = hipMemcpy(g->offsets.ptr, row_offsets, ((size_t)nodes + 1) * sizeof(int32_t),
                                 hipMemcpyHostToDevice);
        if (e == hipSuccess && edges)
            e = hipMemcpy(g->cols.ptr, column_indices, (size_t)edges * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) s = fail(B40C_ERR_HIP, "graph upload failed: %s", hipGetErrorString(e));
    }
    if (s != B40C_OK) {
        delete g;
        return s;
    }
    *out = g;
    return B40C_OK;
}

b40c_status b40c_graph_from_coo(int32_t nodes, int64_t edges, const int32_t* rows, const int32_t* cols,
                                b40c_graph** out) {
    if (!out) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_from_coo: null out");
    b40c_graph* g = new b40c_graph();
    b40c_status s = b40c::build_from_coo(nodes, edges, rows, cols, g);
    if (s != B40C_OK) {
        delete g;
        return s;
    }
    *out = g;
    return B40C_OK;
}

b40c_status b40c_graph_from_coo_unique(int32_t nodes, int64_t edges, const int32_t* rows, const int32_t* cols,
                                       int32_t flags, b40c_graph** out) {
    if (!out) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_from_coo_unique: null out");
    b40c_graph* g = new b40c_graph();
    b40c_status s = b40c::build_from_coo_unique(nodes, edges, rows, cols, flags, g);
    if (s != B40C_OK) {
        delete g;
        return s;
    }
    *out = g;
    return B40C_OK;
}

b40c_status b40c_graph_build_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                                  b40c_graph** out) {
    if (!out) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_build_rmat: null out");
    b40c_graph* g = new b40c_graph();
    b40c_status s = b40c::build_rmat(scale, edge_factor, a, b, c, seed, g);
    if (s != B40C_OK) {
        delete g;
        return s;
    }
    *out = g;
    return B40C_OK;
}

b40c_status b40c_graph_build_lattice3d(int32_t width, uint64_t seed, int32_t* perm, b40c_graph** out) {
    if (!out) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_build_lattice3d: null out");
    b40c_graph* g = new b40c_graph();
    b40c_status s = b40c::build_lattice(width, seed, perm, g);
    if (s != B40C_OK) {
        delete g;
        return s;
    }
    *out = g;
    return B40C_OK;
}

b40c_status b40c_graph_info(const b40c_graph* g, int32_t* nodes, int64_t* edges) {
    if (!g) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_info: null graph");
    if (nodes) *nodes = g->nodes;
    if (edges) *edges = g->edges;
    return B40C_OK;
}

b40c_status b40c_graph_copy_to_host(const b40c_graph* g, int32_t* row_offsets, int32_t* column_indices) {
    if (!g) return fail(B40C_ERR_INVALID_ARG, "b40c_graph_copy_to_host: null graph");
    if (row_offsets)
        B40C_HIP(hipMemcpy(row_offsets, g->offsets.ptr, ((size_t)g->nodes + 1) * sizeof(int32_t),
                           hipMemcpyDeviceToHost));
    if (column_indices && g->edges)
        B40C_HIP(hipMemcpy(column_indices, g->cols.ptr, (size_t)g->edges * sizeof(int32_t), hipMemcpyDeviceToHost));
    return B40C_OK;
}

void b40c_graph_destroy(b40c_graph* g) { b40c::graph_release(g); }

}  // extern "C"
