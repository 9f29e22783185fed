// Batched and served BFS for gfx950: bit-parallel multi-source traversal
// over a packed multi-graph CSR, with per-query label slots and admission at
// level boundaries (continuous batching).  No counterpart in the reference;
// it stands beside the single-source enactors and shares their device
// building blocks (LDS-buffered queues, tiered gather).
//
// Layout:
//  * packed CSR: graphs concatenated, vertex ids global, graph g owns
//    vertices [vbase[g], vbase[g+1]); gid[v] names v's graph.
//  * per-vertex 64-bit masks: visited[v], front[3][v] (rotating by level % 3).
//    Bit b of vertex v in graph g belongs to the query in slot b of graph g
//    (bitq[g*64 + b]), so up to 64 queries per graph share every edge scan.
//  * labels: query q owns depth/pred int32 slots [lbase[q], lbase[q] + V_g),
//    indexed by local vertex id; depth = level + 1 - qstart[q]; pred is a
//    local id (-2 for the source, -1 unreached).
//
// One launch per level: every workgroup derives the direction from device
// counters (top-down push with 64-bit atomic ORs on sparse levels, bottom-up
// pull with per-vertex early exit on dense ones), clears the mask array two
// levels ahead, and records per-graph "new bits" so the host can retire
// finished queries.
#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "bfs_device.hpp"
#include "internal.hpp"

namespace b40c {

struct alignas(256) MsCounters {
    uint32_t qlen[4];
    unsigned long long level_edges[4];
    uint32_t dir[4];
    uint32_t overflow;
    uint32_t corrupt;
    uint32_t ph_base;
    uint32_t pad;
};

struct MsArgs {
    const int32_t* offs;
    const int32_t* cols;
    const int32_t* gid;
    const int32_t* vbase;
    unsigned long long* visited;
    unsigned long long* front[3];
    const unsigned long long* active;  // [G] slots in use per graph
    unsigned long long* newbits;       // [2][G] per-level discoveries by parity
    const int32_t* bitq;               // [G*64] query in each slot (-1 free)
    const long long* lbase;            // [Q] label slot base
    const int32_t* qstart;             // [Q] level the query joined
    int32_t* depth_out;
    int32_t* pred_out;
    int32_t* queue[2];
    uint32_t qcap;
    MsCounters* ctr;
    int32_t n;
    int32_t G;
    int32_t alpha;
    int32_t beta;
    int32_t cta_th;
    int32_t wave_th;
    int32_t dev_base;
    long long active_edges;
    long long active_vertices;
};

__device__ __forceinline__ unsigned long long atomic_or_u64(unsigned long long* p, unsigned long long v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct MsShared {
    LdsQueue<false> q;
    GatherSmem g;
    unsigned long long red[BFS_WAVES];
};

// Labels for every newly reached (query, v) bit; `src` is the discovering
// neighbour (global id) or -1 for a source.
__device__ __forceinline__ void ms_record(const MsArgs& a, int ph, int32_t v, int32_t src, unsigned long long bits) {
    const int32_t g = a.gid[v];
    const int32_t vb = a.vbase[g];
    while (bits) {
        const int b = __ffsll(bits) - 1;
        bits &= bits - 1;
        const int32_t q = a.bitq[g * 64 + b];
        if (q < 0) continue;
        const long long idx = a.lbase[q] + (v - vb);
        a.depth_out[idx] = ph + 1 - a.qstart[q];
        a.pred_out[idx] = src - vb;
    }
}

// Wave-aggregated OR of per-lane discoveries into newbits[slot][g]: lanes
// sharing lane 0's graph are reduced with shuffles, the rest go direct.
__device__ __forceinline__ void ms_note_new(const MsArgs& a, int slot, int32_t g, unsigned long long bits) {
    const int32_t g0 = __shfl(g, 0, kWave);
    const bool same = (g == g0);
    unsigned long long x = same ? bits : 0ull;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x |= __shfl_xor(x, d, kWave);
    if (lane_id() == 0 && x && g0 >= 0) atomic_or_u64(&a.newbits[(size_t)slot * a.G + g0], x);
    if (!same && bits && g >= 0) atomic_or_u64(&a.newbits[(size_t)slot * a.G + g], bits);
}

__device__ __forceinline__ uint32_t ms_direction(const MsArgs& a, int ph, uint32_t n_f) {
    const long long m_f = (long long)a.ctr->level_edges[ph & 3];
    const uint32_t prev = a.ctr->dir[(ph + 3) & 3];
    if (prev == 0) return (m_f * (long long)a.alpha > a.active_edges) ? 1u : 0u;
    return ((long long)n_f * (long long)a.beta < a.active_vertices) ? 0u : 1u;
}

__device__ __forceinline__ void ms_phase_begin(const MsArgs& a, int ph) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st_relaxed(&a.ctr->qlen[(ph + 2) & 3], 0u);
        a.ctr->level_edges[(ph + 2) & 3] = 0ull;
    }
}

__global__ __launch_bounds__(BFS_THREADS) void k_ms_level(MsArgs a, int ph) {
    __shared__ MsShared s;
    if (a.dev_base) ph += (int)ld_relaxed(&a.ctr->ph_base);
    uint32_t n_f = ld_relaxed(&a.ctr->qlen[ph & 3]);
    if (n_f > a.qcap) n_f = 0;
    const int nslot = (ph + 1) & 1;
    // Clear next level's new-bit slot (read by the host after this level) and
    // the mask array two levels ahead.
    for (int32_t i = blockIdx.x * BFS_THREADS + threadIdx.x; i < a.G; i += gridDim.x * BFS_THREADS)
        a.newbits[(size_t)nslot * a.G + i] = 0ull;
    unsigned long long* clear = a.front[(ph + 2) % 3];
    for (int32_t v = blockIdx.x * BFS_THREADS + threadIdx.x; v < a.n; v += gridDim.x * BFS_THREADS) clear[v] = 0ull;
    if (n_f == 0) {
        ms_phase_begin(a, ph);
        return;
    }
    const uint32_t dir = ms_direction(a, ph, n_f);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctr->dir[ph & 3] = dir;
    if (threadIdx.x == 0) {
        s.q.count = 0;
    }
    gather_init(s.g);
    __syncthreads();
    ms_phase_begin(a, ph);
    GlobalQueue out;
    out.keys = a.queue[(ph + 1) & 1];
    out.vals = nullptr;
    out.len = &a.ctr->qlen[(ph + 1) & 3];
    out.cap = a.qcap;
    out.overflow = &a.ctr->overflow;
    const unsigned long long* cur = a.front[ph % 3];
    unsigned long long* next = a.front[(ph + 1) % 3];
    const int slot = ph & 1;  // newbits slot written by this level
    unsigned long long eacc = 0;

    if (dir == 0) {
        // Top-down push over the frontier queue.
        const int32_t* qin = a.queue[ph & 1];
        for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_f; base += gridDim.x * BFS_THREADS) {
            const uint32_t i = base + threadIdx.x;
            int32_t u = -1, start = 0, deg = 0;
            if (i < n_f) {
                u = qin[i];
                if ((uint32_t)u < (uint32_t)a.n) {
                    start = a.offs[u];
                    deg = a.offs[u + 1] - start;
                } else {
                    st_relaxed(&a.ctr->corrupt, 1u);
                    u = -1;
                }
            }
            auto sink = [&](bool valid, int32_t v, int32_t src) {
                unsigned long long nb = 0;
                int32_t g = -1;
                if (valid) {
                    const unsigned long long w = cur[src] & ~a.visited[v];
                    if (w) {
                        const unsigned long long old = atomic_or_u64(&a.visited[v], w);
                        nb = w & ~old;
                    }
                }
                bool enq = false;
                if (nb) {
                    g = a.gid[v];
                    ms_record(a, ph, v, src, nb);
                    enq = atomic_or_u64(&next[v], nb) == 0ull;
                    if (enq) eacc += (unsigned long long)(a.offs[v + 1] - a.offs[v]);
                }
                ms_note_new(a, slot, g, nb);
                lq_append(s.q, out, enq, v, 0);
            };
            tiered_gather(a.cols, u, start, deg, a.cta_th, a.wave_th, s.g, sink);
            lq_flush(s.q, out, LQ_CAP / 2);
        }
    } else {
        // Bottom-up pull over every vertex with bits still wanted.
        for (int32_t v = blockIdx.x * BFS_THREADS + threadIdx.x; v - (int32_t)threadIdx.x < a.n;
             v += gridDim.x * BFS_THREADS) {
            unsigned long long got = 0;
            int32_t g = -1;
            if (v < a.n) {
                g = a.gid[v];
                const unsigned long long vis = a.visited[v];
                const unsigned long long want = a.active[g] & ~vis;
                if (want) {
                    const int32_t e1 = a.offs[v + 1];
                    for (int32_t e = a.offs[v]; e < e1; ++e) {
                        const int32_t u = a.cols[e];
                        const unsigned long long f = cur[u] & want & ~got;
                        if (f) {
                            got |= f;
                            ms_record(a, ph, v, u, f);
                            if (got == want) break;
                        }
                    }
                    if (got) {
                        a.visited[v] = vis | got;
                        next[v] = got;
                        eacc += (unsigned long long)(e1 - a.offs[v]);
                    }
                }
            }
            ms_note_new(a, slot, g, got);
            lq_append(s.q, out, got != 0ull, v, 0);
        }
    }
    lq_flush(s.q, out, 1);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) eacc += __shfl_xor(eacc, d, kWave);
    __syncthreads();
    if (lane_id() == 0) s.red[threadIdx.x / kWave] = eacc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < BFS_WAVES; ++w) t += s.red[w];
        if (t) atomic_add_u64(&a.ctr->level_edges[(ph + 1) & 3], t);
    }
}

__global__ void k_ms_advance(MsArgs a, int steps) {
    if (threadIdx.x == 0 && blockIdx.x == 0) a.ctr->ph_base += (uint32_t)steps;
}

// Per admitted query: graph slot, label base and start level.
struct MsSeed {
    int32_t q;
    int32_t src;  // global id
    int32_t bit;
    int32_t qstart;
    long long lbase;
};

// Registers newly admitted queries (slot table, active mask, label base,
// start level) and seeds them into level ph: source bit in visited and the
// current mask array, source labels, and the source appended to the current
// frontier queue if it was not already on it.  The slot table and masks are
// only written here and in k_ms_release, never copied from the host.
__global__ void k_ms_seed(MsArgs a, int ph, const MsSeed* seeds, int count, int32_t* bitq,
                          unsigned long long* active, long long* lbase, int32_t* qstart) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const MsSeed sd = seeds[i];
    const int32_t q = sd.q, v = sd.src;
    if ((uint32_t)v >= (uint32_t)a.n) {
        st_relaxed(&a.ctr->corrupt, 1u);
        return;
    }
    const unsigned long long bit = 1ull << sd.bit;
    const int32_t g = a.gid[v];
    bitq[g * 64 + sd.bit] = q;
    lbase[q] = sd.lbase;
    qstart[q] = sd.qstart;
    atomic_or_u64(&active[g], bit);
    atomic_or_u64(&a.visited[v], bit);
    const long long idx = sd.lbase + (v - a.vbase[g]);
    a.depth_out[idx] = 0;
    a.pred_out[idx] = -2;
    const unsigned long long old = atomic_or_u64(&a.front[ph % 3][v], bit);
    if (old == 0ull) {
        const uint32_t pos = atomic_add_u32(&a.ctr->qlen[ph & 3], 1u);
        if (pos < a.qcap)
            a.queue[ph & 1][pos] = v;
        else
            st_relaxed(&a.ctr->overflow, 1u);
        atomic_add_u64(&a.ctr->level_edges[ph & 3], (unsigned long long)(a.offs[v + 1] - a.offs[v]));
    }
}

// Releases graph g's slot `bit`: slot table entry, active bit, and the bit in
// visited over the graph's vertex range [v0, v1).
__global__ void k_ms_release(unsigned long long* visited, int32_t v0, int32_t v1, int32_t g, int32_t bit,
                             int32_t* bitq, unsigned long long* active) {
    const unsigned long long keep = ~(1ull << bit);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        bitq[g * 64 + bit] = -1;
        __hip_atomic_fetch_and(&active[g], keep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (int32_t v = v0 + blockIdx.x * blockDim.x + threadIdx.x; v < v1; v += gridDim.x * blockDim.x)
        visited[v] &= keep;
}

__global__ void k_fill_i32(int32_t* p, int32_t value, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

// Packs several CSRs into one: vertex ids shifted by the graph's base.
__global__ void k_pack_cols(const int32_t* cols, int64_t m, int32_t shift, int32_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = cols[i] + shift;
}
__global__ void k_pack_offs(const int32_t* offs, int32_t n, int32_t eshift, int32_t gidx, int32_t* out,
                            int32_t* gid) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
        out[i] = offs[i] + eshift;
        if (i < n) gid[i] = gidx;
    }
}

}  // namespace b40c

// ===========================================================================
// Host side.
// ===========================================================================

struct b40c_multigraph {
    b40c_graph* packed = nullptr;
    int32_t ngraphs = 0;
    std::vector<int32_t> vbase;  // ngraphs + 1
    std::vector<int64_t> ebase;  // ngraphs + 1
    b40c::DeviceBuffer d_vbase, d_gid;
};

namespace b40c {
namespace {

struct MsQuery {
    int32_t graph = -1;
    int32_t src = -1;  // local id
    int32_t bit = -1;
    int64_t lbase = -1;
    int32_t qstart = 0;
    int32_t state = 0;  // 0 waiting, 1 running, 2 done, 3 released
    int32_t levels = 0;
};

// First-fit allocator over the label arena (units: int32 elements).
struct ArenaAlloc {
    std::map<int64_t, int64_t> free_;  // start -> length
    void init(int64_t n) {
        free_.clear();
        if (n > 0) free_[0] = n;
    }
    int64_t alloc(int64_t n) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                const int64_t s = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[s + n] = len - n;
                return s;
            }
        }
        return -1;
    }
    void release(int64_t s, int64_t n) {
        auto it = free_.emplace(s, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}  // namespace
}  // namespace b40c

struct b40c_msbfs {
    b40c_multigraph* mg = nullptr;
    hipStream_t stream = nullptr;
    int32_t n = 0, G = 0;
    int64_t arena_elems = 0;  // per label array (depth and pred each)
    b40c::DeviceBuffer visited, front[3], active, newbits, bitq, lbase, qstart, depth, pred, queue[2], counters;
    b40c::DeviceBuffer seed_buf;
    b40c::MsSeed* h_seeds = nullptr;  // pinned staging
    int h_seed_cap = 0;
    uint32_t qcap = 0;
    b40c::ArenaAlloc arena;
    std::vector<b40c::MsQuery> queries;
    std::deque<int32_t> waiting;
    std::vector<unsigned long long> h_active;  // [G]
    std::vector<int32_t> running;              // query ids in flight
    std::vector<int32_t> completed;            // finished, not yet reported
    int32_t ph = 0;
    int32_t alpha = 14, beta = 24, grid = 1024;
    b40c::MsCounters* h_ctr = nullptr;
    std::vector<unsigned long long> h_newbits;
    int32_t qmax = 0;
};

namespace b40c {
namespace {

MsArgs ms_args(b40c_msbfs* m) {
    MsArgs a{};
    a.offs = m->mg->packed->d_offsets();
    a.cols = m->mg->packed->d_cols();
    a.gid = m->mg->d_gid.as<int32_t>();
    a.vbase = m->mg->d_vbase.as<int32_t>();
    a.visited = m->visited.as<unsigned long long>();
    for (int i = 0; i < 3; ++i) a.front[i] = m->front[i].as<unsigned long long>();
    a.active = m->active.as<unsigned long long>();
    a.newbits = m->newbits.as<unsigned long long>();
    a.bitq = m->bitq.as<int32_t>();
    a.lbase = m->lbase.as<long long>();
    a.qstart = m->qstart.as<int32_t>();
    a.depth_out = m->depth.as<int32_t>();
    a.pred_out = m->pred.as<int32_t>();
    a.queue[0] = m->queue[0].as<int32_t>();
    a.queue[1] = m->queue[1].as<int32_t>();
    a.qcap = m->qcap;
    a.ctr = m->counters.as<MsCounters>();
    a.n = m->n;
    a.G = m->G;
    a.alpha = m->alpha;
    a.beta = m->beta;
    a.cta_th = 256;
    a.wave_th = 64;
    a.dev_base = 0;
    int64_t ae = 0, av = 0;
    for (int g = 0; g < m->G; ++g)
        if (m->h_active[g]) {
            ae += m->mg->ebase[g + 1] - m->mg->ebase[g];
            av += m->mg->vbase[g + 1] - m->mg->vbase[g];
        }
    a.active_edges = std::max<int64_t>(ae, 1);
    a.active_vertices = std::max<int64_t>(av, 1);
    return a;
}

unsigned grid_for_n(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// Admits waiting queries while a slot of their graph and label memory are free.
b40c_status ms_admit(b40c_msbfs* m) {
    std::vector<MsSeed> seeds;
    std::deque<int32_t> still;
    while (!m->waiting.empty()) {
        const int32_t q = m->waiting.front();
        m->waiting.pop_front();
        MsQuery& Q = m->queries[q];
        const int32_t g = Q.graph;
        const int64_t vg = m->mg->vbase[g + 1] - m->mg->vbase[g];
        const unsigned long long used = m->h_active[g];
        if (used == ~0ull) {
            still.push_back(q);
            continue;
        }
        const int64_t lb = m->arena.alloc(vg);
        if (lb < 0) {
            still.push_back(q);
            continue;
        }
        const int b = __builtin_ctzll(~used);
        Q.bit = b;
        Q.lbase = lb;
        Q.qstart = m->ph;
        Q.state = 1;
        m->h_active[g] |= 1ull << b;
        // Label slot := -1 for depth and pred.
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for_n(vg)), dim3(256), 0, m->stream, m->depth.as<int32_t>() + lb, -1,
                           (long long)vg);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for_n(vg)), dim3(256), 0, m->stream, m->pred.as<int32_t>() + lb, -1,
                           (long long)vg);
        seeds.push_back(MsSeed{q, m->mg->vbase[g] + Q.src, b, m->ph, (long long)lb});
        m->running.push_back(q);
    }
    m->waiting.swap(still);
    if (seeds.empty()) return B40C_OK;
    const int cnt = (int)seeds.size();
    // Pinned staging: the previous step synchronized the stream, so the
    // buffer is free to overwrite.
    if (m->h_seed_cap < cnt) {
        if (m->h_seeds) (void)hipHostFree(m->h_seeds);
        m->h_seeds = nullptr;
        m->h_seed_cap = 0;
        B40C_HIP(hipHostMalloc((void**)&m->h_seeds, sizeof(MsSeed) * (size_t)std::max(cnt, 1024), 0));
        m->h_seed_cap = std::max(cnt, 1024);
    }
    std::memcpy(m->h_seeds, seeds.data(), sizeof(MsSeed) * cnt);
    B40C_TRY(m->seed_buf.reserve(sizeof(MsSeed) * (size_t)m->h_seed_cap));
    B40C_HIP(hipMemcpyAsync(m->seed_buf.ptr, m->h_seeds, sizeof(MsSeed) * cnt, hipMemcpyHostToDevice, m->stream));
    const MsArgs a = ms_args(m);
    hipLaunchKernelGGL(k_ms_seed, dim3((cnt + 255) / 256), dim3(256), 0, m->stream, a, m->ph,
                       m->seed_buf.as<MsSeed>(), cnt, m->bitq.as<int32_t>(), m->active.as<unsigned long long>(),
                       m->lbase.as<long long>(), m->qstart.as<int32_t>());
    B40C_HIP(hipGetLastError());
    return B40C_OK;
}

// Retires queries whose slot found nothing new in the level just run.
b40c_status ms_retire(b40c_msbfs* m) {
    const int slot = m->ph & 1;  // the level that just ran wrote this slot
    std::vector<int32_t> keep;
    for (int32_t q : m->running) {
        MsQuery& Q = m->queries[q];
        const unsigned long long bit = 1ull << Q.bit;
        Q.levels = m->ph + 1 - Q.qstart;
        if (m->h_newbits[(size_t)slot * m->G + Q.graph] & bit) {
            keep.push_back(q);
            continue;
        }
        Q.state = 2;
        m->completed.push_back(q);
        m->h_active[Q.graph] &= ~bit;
        const int32_t v0 = m->mg->vbase[Q.graph], v1 = m->mg->vbase[Q.graph + 1];
        hipLaunchKernelGGL(k_ms_release, dim3(grid_for_n(v1 - v0)), dim3(256), 0, m->stream,
                           m->visited.as<unsigned long long>(), v0, v1, Q.graph, Q.bit, m->bitq.as<int32_t>(),
                           m->active.as<unsigned long long>());
        B40C_HIP(hipGetLastError());
    }
    m->running.swap(keep);
    return B40C_OK;
}

}  // namespace
}  // namespace b40c

using namespace b40c;

extern "C" {

b40c_status b40c_multigraph_pack(b40c_graph* const* graphs, int32_t ngraphs, b40c_multigraph** out) {
    if (!graphs || ngraphs < 1 || !out) return fail(B40C_ERR_INVALID_ARG, "b40c_multigraph_pack: bad arguments");
    b40c_multigraph* mg = new b40c_multigraph();
    mg->ngraphs = ngraphs;
    mg->vbase.assign(ngraphs + 1, 0);
    mg->ebase.assign(ngraphs + 1, 0);
    for (int g = 0; g < ngraphs; ++g) {
        if (!graphs[g]) {
            delete mg;
            return fail(B40C_ERR_INVALID_ARG, "b40c_multigraph_pack: graph %d is null", g);
        }
        mg->vbase[g + 1] = mg->vbase[g] + graphs[g]->nodes;
        mg->ebase[g + 1] = mg->ebase[g] + graphs[g]->edges;
        if ((int64_t)mg->vbase[g + 1] > 0x7FFFFFFFll || mg->ebase[g + 1] > 0x7FFFFFFFll) {
            delete mg;
            return fail(B40C_ERR_UNSUPPORTED, "b40c_multigraph_pack: packed graph exceeds int32 ids/offsets");
        }
    }
    const int32_t N = mg->vbase[ngraphs];
    const int64_t M = mg->ebase[ngraphs];
    b40c_graph* P = new b40c_graph();
    P->nodes = N;
    P->edges = M;
    b40c_status s = P->offsets.alloc(((size_t)N + 1) * 4);
    if (s == B40C_OK) s = P->cols.alloc((size_t)std::max<int64_t>(M, 1) * 4);
    if (s == B40C_OK) s = mg->d_gid.alloc((size_t)N * 4);
    if (s == B40C_OK) s = mg->d_vbase.alloc(((size_t)ngraphs + 1) * 4);
    if (s != B40C_OK) {
        delete P;
        delete mg;
        return s;
    }
    for (int g = 0; g < ngraphs; ++g) {
        const b40c_graph* src = graphs[g];
        if (src->edges)
            hipLaunchKernelGGL(k_pack_cols, dim3(grid_for_n(src->edges)), dim3(256), 0, nullptr, src->d_cols(),
                               src->edges, mg->vbase[g], P->cols.as<int32_t>() + mg->ebase[g]);
        hipLaunchKernelGGL(k_pack_offs, dim3(grid_for_n(src->nodes + 1)), dim3(256), 0, nullptr, src->d_offsets(),
                           src->nodes, (int32_t)mg->ebase[g], g, P->offsets.as<int32_t>() + mg->vbase[g],
                           mg->d_gid.as<int32_t>() + mg->vbase[g]);
    }
    bool all_sym = true;
    for (int g = 0; g < ngraphs; ++g) all_sym = all_sym && graphs[g]->symmetric.load() == 1;
    P->symmetric.store(all_sym ? 1 : -1);
    hipError_t e = hipMemcpy(mg->d_vbase.ptr, mg->vbase.data(), mg->vbase.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        delete P;
        delete mg;
        return fail(B40C_ERR_HIP, "b40c_multigraph_pack: %s", hipGetErrorString(e));
    }
    mg->packed = P;
    *out = mg;
    return B40C_OK;
}

b40c_status b40c_multigraph_info(const b40c_multigraph* mg, int32_t* ngraphs, int32_t* vbase, int64_t* ebase) {
    if (!mg) return fail(B40C_ERR_INVALID_ARG, "b40c_multigraph_info: null");
    if (ngraphs) *ngraphs = mg->ngraphs;
    if (vbase) std::memcpy(vbase, mg->vbase.data(), mg->vbase.size() * 4);
    if (ebase) std::memcpy(ebase, mg->ebase.data(), mg->ebase.size() * 8);
    return B40C_OK;
}

void b40c_multigraph_destroy(b40c_multigraph* mg) {
    if (!mg) return;
    graph_release(mg->packed);
    delete mg;
}

b40c_status b40c_msbfs_create(b40c_multigraph* mg, int64_t label_budget_bytes, int32_t max_queries,
                              b40c_msbfs** out) {
    if (!mg || !out || label_budget_bytes < 8 || max_queries < 1)
        return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_create: bad arguments");
    bool sym = false;
    B40C_TRY(graph_symmetric(mg->packed, sym));
    if (!sym)
        return fail(B40C_ERR_UNSUPPORTED, "b40c_msbfs_create: the batched engine needs symmetric graphs "
                                          "(bottom-up steps pull from neighbours)");
    b40c_msbfs* m = new b40c_msbfs();
    m->mg = mg;
    m->n = mg->packed->nodes;
    m->G = mg->ngraphs;
    m->arena_elems = label_budget_bytes / 8;
    m->qmax = max_queries;
    m->qcap = (uint32_t)m->n;
    DeviceInfo info;
    b40c_status s = device_info(info);
    m->grid = info.num_cus * 4;
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&m->h_ctr, sizeof(MsCounters), hipHostMallocDefault);
    if (e != hipSuccess) s = fail(B40C_ERR_HIP, "b40c_msbfs_create: %s", hipGetErrorString(e));
    const size_t N = (size_t)m->n;
    if (s == B40C_OK) s = m->visited.alloc(N * 8);
    for (int i = 0; i < 3 && s == B40C_OK; ++i) s = m->front[i].alloc(N * 8);
    if (s == B40C_OK) s = m->active.alloc((size_t)m->G * 8);
    if (s == B40C_OK) s = m->newbits.alloc((size_t)m->G * 16);
    if (s == B40C_OK) s = m->bitq.alloc((size_t)m->G * 64 * 4);
    if (s == B40C_OK) s = m->lbase.alloc((size_t)max_queries * 8);
    if (s == B40C_OK) s = m->qstart.alloc((size_t)max_queries * 4);
    if (s == B40C_OK) s = m->depth.alloc((size_t)m->arena_elems * 4);
    if (s == B40C_OK) s = m->pred.alloc((size_t)m->arena_elems * 4);
    if (s == B40C_OK) s = m->queue[0].alloc(N * 4);
    if (s == B40C_OK) s = m->queue[1].alloc(N * 4);
    if (s == B40C_OK) s = m->counters.alloc(sizeof(MsCounters));
    if (s != B40C_OK) {
        b40c_msbfs_destroy(m);
        return s;
    }
    m->arena.init(m->arena_elems);
    m->h_active.assign(m->G, 0ull);
    m->h_newbits.assign((size_t)m->G * 2, 0ull);
    B40C_HIP(hipMemsetAsync(m->visited.ptr, 0, m->visited.bytes, m->stream));
    for (int i = 0; i < 3; ++i) B40C_HIP(hipMemsetAsync(m->front[i].ptr, 0, m->front[i].bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->newbits.ptr, 0, m->newbits.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->counters.ptr, 0, sizeof(MsCounters), m->stream));
    B40C_HIP(hipMemsetAsync(m->bitq.ptr, 0xFF, m->bitq.bytes, m->stream));  // all slots free (-1)
    B40C_HIP(hipMemsetAsync(m->active.ptr, 0, m->active.bytes, m->stream));
    B40C_HIP(hipStreamSynchronize(m->stream));
    *out = m;
    return B40C_OK;
}

b40c_status b40c_msbfs_set_tuning(b40c_msbfs* m, int32_t alpha, int32_t beta, int32_t grid) {
    if (!m) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_set_tuning: null");
    if (alpha > 0) m->alpha = alpha;
    if (beta > 0) m->beta = beta;
    if (grid > 0) m->grid = grid;
    return B40C_OK;
}

b40c_status b40c_msbfs_submit(b40c_msbfs* m, int32_t graph, int32_t src, int32_t* query_id) {
    if (!m || !query_id) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_submit: null");
    if (graph < 0 || graph >= m->G) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_submit: graph %d out of range", graph);
    const int32_t vg = m->mg->vbase[graph + 1] - m->mg->vbase[graph];
    if (src < 0 || src >= vg) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_submit: source %d out of range", src);
    if ((int64_t)vg > m->arena_elems)
        return fail(B40C_ERR_ADMISSION, "b40c_msbfs_submit: labels of graph %d (%d vertices) exceed the budget", graph,
                    vg);
    if ((int32_t)m->queries.size() >= m->qmax)
        return fail(B40C_ERR_ADMISSION, "b40c_msbfs_submit: more than %d queries", m->qmax);
    MsQuery Q;
    Q.graph = graph;
    Q.src = src;
    m->queries.push_back(Q);
    const int32_t q = (int32_t)m->queries.size() - 1;
    m->waiting.push_back(q);
    *query_id = q;
    return B40C_OK;
}

// One level boundary: admit, run one level for every query in flight, retire.
b40c_status b40c_msbfs_step(b40c_msbfs* m, int32_t* running, int32_t* waiting) {
    if (!m) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_step: null");
    B40C_TRY(ms_admit(m));
    if (!m->running.empty()) {
        const MsArgs a = ms_args(m);
        hipLaunchKernelGGL(k_ms_level, dim3(m->grid), dim3(BFS_THREADS), 0, m->stream, a, m->ph);
        B40C_HIP(hipGetLastError());
        B40C_HIP(hipMemcpyAsync(m->h_ctr, m->counters.ptr, sizeof(MsCounters), hipMemcpyDeviceToHost, m->stream));
        B40C_HIP(hipMemcpyAsync(m->h_newbits.data(), m->newbits.ptr, m->h_newbits.size() * 8, hipMemcpyDeviceToHost,
                                m->stream));
        B40C_HIP(hipStreamSynchronize(m->stream));
        if (m->h_ctr->overflow || m->h_ctr->corrupt)
            return fail(B40C_ERR_HIP, "batched traversal: frontier overflow or corrupt frontier");
        B40C_TRY(ms_retire(m));
        m->ph += 1;
    }
    if (running) *running = (int32_t)m->running.size();
    if (waiting) *waiting = (int32_t)m->waiting.size();
    return B40C_OK;
}

// Runs level steps until every submitted query has finished.
b40c_status b40c_msbfs_run(b40c_msbfs* m) {
    if (!m) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_run: null");
    int32_t r = 0, w = 0;
    do {
        B40C_TRY(b40c_msbfs_step(m, &r, &w));
        if (r == 0 && w > 0)
            return fail(B40C_ERR_ADMISSION, "b40c_msbfs_run: %d queries can never be admitted", w);
    } while (r > 0 || w > 0);
    return B40C_OK;
}

// Completed queries not yet returned by an earlier poll.
b40c_status b40c_msbfs_poll(b40c_msbfs* m, int32_t* ids, int32_t max_ids, int32_t* count) {
    if (!m || !count || (max_ids > 0 && !ids)) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_poll: bad arguments");
    const int32_t k = std::min<int32_t>(max_ids, (int32_t)m->completed.size());
    for (int32_t i = 0; i < k; ++i) ids[i] = m->completed[i];
    m->completed.erase(m->completed.begin(), m->completed.begin() + k);
    *count = k;
    return B40C_OK;
}

// Copies a finished query's labels (local ids; either pointer may be NULL),
// reports its search depth, and frees its label slot.
b40c_status b40c_msbfs_result(b40c_msbfs* m, int32_t q, int32_t* depth, int32_t* pred, int32_t* search_depth) {
    if (!m || q < 0 || q >= (int32_t)m->queries.size()) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_result: bad id");
    MsQuery& Q = m->queries[q];
    if (Q.state != 2) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_result: query %d is not finished", q);
    const int64_t vg = m->mg->vbase[Q.graph + 1] - m->mg->vbase[Q.graph];
    if (depth)
        B40C_HIP(hipMemcpyAsync(depth, m->depth.as<int32_t>() + Q.lbase, vg * 4, hipMemcpyDeviceToHost, m->stream));
    if (pred) B40C_HIP(hipMemcpyAsync(pred, m->pred.as<int32_t>() + Q.lbase, vg * 4, hipMemcpyDeviceToHost, m->stream));
    B40C_HIP(hipStreamSynchronize(m->stream));
    if (search_depth) *search_depth = Q.levels;
    m->arena.release(Q.lbase, vg);
    Q.state = 3;
    return B40C_OK;
}

void* b40c_msbfs_stream(b40c_msbfs* m) { return m ? (void*)m->stream : nullptr; }

void b40c_msbfs_destroy(b40c_msbfs* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->visited.release();
    for (int i = 0; i < 3; ++i) m->front[i].release();
    m->active.release();
    m->newbits.release();
    m->bitq.release();
    m->lbase.release();
    m->qstart.release();
    m->depth.release();
    m->pred.release();
    m->queue[0].release();
    m->queue[1].release();
    m->counters.release();
    m->seed_buf.release();
    if (m->h_seeds) (void)hipHostFree(m->h_seeds);
    if (m->h_ctr) (void)hipHostFree(m->h_ctr);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

}  // extern "C"
