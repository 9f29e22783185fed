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
// Direction is chosen per query, not per level: queries arrive at different
// times, so one level mixes young queries (a handful of frontier vertices)
// with mature ones (frontiers covering much of their graph).  Each level
// runs these launches:
//  * k_ms_push + k_ms_heavy: top-down push over the frontier queue for the
//    bits in top-down mode (64-bit atomic ORs); long lists go to a
//    load-balanced grid tier.
//  * k_ms_pull_narrow + k_ms_pull: bottom-up pull over the vertices of
//    graphs with bits in bottom-up mode (each vertex scans its neighbours
//    until every wanted bit is found; unresolved lists are scanned by 8-lane
//    groups), in a narrow form for graphs with a few such bits and a wide
//    form (staged, coalesced label stores) for graphs with many.  A young
//    query's bit is never in a bottom-up want set, so the pull's early exit
//    keeps working while queries of every age share it.
//  * k_ms_count: per (graph, slot) count and degree sum of the vertices
//    discovered, by a 64x64 bit transpose across each wave; also the next
//    level's one-bit-per-vertex frontier bitmap.
//  * k_ms_direction: per slot, Beamer's rule (Beamer, Asanovic, Patterson,
//    SC 2012) on the slot's own counts -> next level's bottom-up masks, and
//    the per-graph "found anything" mask the host uses to retire queries.
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
    uint32_t any_bu[2];  // some slot is in bottom-up mode (by level parity)
    uint32_t overflow;
    uint32_t corrupt;
    uint32_t bar_abort;  // persistent form: a grid-barrier wait timed out
    uint32_t pad[7];
    unsigned long long heavy_alloc[2];  // (entries << 32 | chunks) deferred to the grid tier, by parity
    unsigned long long bu_tiles[2];     // (graphs << 32 | tiles) listed for the wide pull, by parity
    unsigned long long bu_tiles_n[2];   // the same for the narrow pull
    unsigned long long bar_release;     // persistent form: grid-barrier release word (epoch | payload)
};

struct MsArgs {
    const int32_t* offs;
    const int32_t* cols;
    const int32_t* gid;
    const int32_t* vbase;
    const long long* ebase;            // [G+1]
    unsigned long long* visited;
    unsigned long long* front[3];
    unsigned long long* active;        // [G] slots in use per graph
    unsigned long long* bu;            // [G] slots in bottom-up mode this level
    uint32_t* cnt_n;                   // [G*64] vertices discovered in the last level, per slot
    unsigned long long* cnt_m;         // [G*64] their degree sum
    uint32_t* nf_prev;                 // [G*64] the level before's count
    unsigned long long* m_seen;        // [G*64] degree sum of everything discovered so far
    int32_t* bitq;                     // [G*64] query in each slot (-1 free)
    long long* lbase;                  // [Q] label slot base
    int32_t* qstart;                   // [Q] level the query joined
    int32_t* depth_out;
    int32_t* pred_out;
    int32_t* queue[3];                 // frontier queues, rotating by level % 3 (front[] pairs with queue[])
    uint32_t* fbm[2];                  // [ceil(n/32)] by level parity: vertices carrying any bit in front[]
    int2* bu_list;                     // [2 classes][2 parities][G]: (graph, first tile) of graphs with bottom-up slots
    MsCounters* hctr;                  // device address of the pinned host mirror (overflow, corrupt)
    int32_t* hdone;                    // [Q] device address of pinned host words: levels of a finished query
    int32_t* dph;                      // captured level loops: phase of the replay's first level (else null)
    uint32_t* bar_arrive;              // persistent form: grid-barrier arrival word per workgroup
    int2* heavy;                       // grid-tier entries (vertex, first chunk)
    int32_t* chunk_map;                // chunk -> entry
    uint32_t qcap;
    uint32_t heavy_cap;
    uint32_t chunk_cap;
    MsCounters* ctr;
    int32_t n;
    int32_t G;
    int32_t alpha;
    int32_t beta;
    int32_t cta_th;
    int32_t wave_th;
    int32_t grid_th;
};

__device__ __forceinline__ unsigned long long atomic_or_u64(unsigned long long* p, unsigned long long v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_and_u64(unsigned long long* p, unsigned long long v) {
    (void)__hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A level kernel's phase: the launch argument itself, or, in a captured
// level loop, the argument (the level's index within the replay) plus the
// replay's first phase, which k_ms_advance moves on at the end of a replay.
__device__ __forceinline__ int ms_phase(const MsArgs& a, int ph) {
    return a.dph ? (int)ld_relaxed((const uint32_t*)a.dph) + ph : ph;
}

// 64x64 bit-matrix transpose across a wave: on return, bit i of lane b's
// word is bit b of lane i's input.  Six block-swap stages of width 32..1.
__device__ __forceinline__ unsigned long long wave_transpose64(unsigned long long x) {
    const int lane = lane_id();
    const unsigned long long lo[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                      0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int s = 32 >> k;
        const unsigned long long p = __shfl_xor(x, s, kWave);
        if (lane & s)
            x = (x & ~lo[k]) | ((p & ~lo[k]) >> s);
        else
            x = (x & lo[k]) | ((p & lo[k]) << s);
    }
    return x;
}

// A per-level pass over the vertices of a frontier queue replaces a pass over
// all n vertices while the queue is shorter than n / MS_SPARSE_DIV.
constexpr int MS_SPARSE_DIV = 16;
// The pull walks the tiles of at most this many graphs through an LDS table
// (more: every tile of the pack).
constexpr int MS_TILE_TABLE = 512;
constexpr int MS_TILE = 64;          // vertices per wide-pull tile (one per lane of a wave)
constexpr int MS_TILE_N = BFS_THREADS;  // vertices per narrow-pull tile (one per thread of a workgroup)
// A graph pulls in the wide form when more than this many of its slots are
// in bottom-up mode, else in the narrow form.
constexpr int MS_WIDE_BITS = 8;
constexpr int kMsMaxFusedGrid = 4096;  // grid-barrier arrival words of the persistent form

struct MsShared {
    LdsKeyQueue q;
    GatherSmem g;
    unsigned long long red[BFS_WAVES];
};

// Labels for every newly reached (query, v) bit; `src` is the discovering
// neighbour (global id) or -1 for a source.  One bit at a time: unrolling
// the slot lookups four bits at a time measured the same and doubled the
// pull's code.
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

__device__ __forceinline__ GlobalQueue ms_out_queue(const MsArgs& a, int ph) {
    GlobalQueue out;
    out.keys = a.queue[(ph + 1) % 3];
    out.vals = nullptr;
    out.len = &a.ctr->qlen[(ph + 1) & 3];
    out.cap = a.qcap;
    out.overflow = &a.ctr->overflow;
    return out;
}

// Top-down edge (src -> v) for the bits of src in top-down mode: exact
// claim by atomic OR on visited[v], labels, and the first OR into next[v]
// enqueues v.  Wave-collective (the LDS queue append).
__device__ __forceinline__ void ms_push_edge(const MsArgs& a, int ph, MsShared& s, const GlobalQueue& out,
                                             bool valid, int32_t v, int32_t src) {
    const unsigned long long* cur = a.front[ph % 3];
    unsigned long long nb = 0;
    if (valid) {
        const unsigned long long w = cur[src] & ~a.bu[a.gid[src]] & ~a.visited[v];
        if (w) {
            const unsigned long long old = atomic_or_u64(&a.visited[v], w);
            nb = w & ~old;
        }
    }
    bool enq = false;
    if (nb) {
        ms_record(a, ph, v, src, nb);
        enq = atomic_or_u64(&a.front[(ph + 1) % 3][v], nb) == 0ull;
    }
    lq_append<false>(s.q, out, enq, v, 0);
}

// Grid tier: a list of at least grid_th edges is cut into HEAVY_CHUNK-edge
// chunks taken by all workgroups of k_ms_heavy.  One 64-bit atomic per wave
// allocates (entries << 32 | chunks); each chunk points at its entry.
__device__ __forceinline__ void ms_defer_heavy(const MsArgs& a, int ph, int32_t v, int32_t& deg) {
    const bool heavy = deg >= a.grid_th;
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
            a.heavy[entry] = make_int2(v, (int)c0);
            for (uint32_t c = 0; c < nch; ++c) a.chunk_map[c0 + c] = (int32_t)entry;
        } else {
            st_relaxed(&a.ctr->overflow, 1u);
        }
        deg = 0;
    }
}

// Top-down push of level ph (lists below the grid tier's threshold).
// Vertices that gain bits are appended once (the first OR into next[v]
// enqueues) to the next frontier queue.
__device__ void ms_push_body(const MsArgs& a, int ph, MsShared& s) {
    uint32_t n_f = ld_relaxed(&a.ctr->qlen[ph & 3]);
    if (n_f > a.qcap) n_f = 0;
    // Clear the mask array of level ph - 1 (it becomes level ph + 2's next
    // array).  Its nonzero entries are exactly the vertices on level ph - 1's
    // queue (every first OR enqueues), so a short queue is cleared entry by
    // entry and a long one by a full pass.  Also the queue length two levels
    // ahead, the next level's grid-tier allocation and its pull tile list.
    unsigned long long* clear = a.front[(ph + 2) % 3];
    const uint32_t n_old = ld_relaxed(&a.ctr->qlen[(ph + 3) & 3]);
    const int32_t gtid = blockIdx.x * BFS_THREADS + threadIdx.x, gstride = gridDim.x * BFS_THREADS;
    if (n_old <= a.qcap && (int64_t)n_old * MS_SPARSE_DIV < (int64_t)a.n) {
        const int32_t* qold = a.queue[(ph + 2) % 3];
        for (int32_t i = gtid; i < (int32_t)n_old; i += gstride) {
            const int32_t v = qold[i];
            if ((uint32_t)v < (uint32_t)a.n) clear[v] = 0ull;
        }
    } else {
        for (int32_t v = gtid; v < a.n; v += gstride) clear[v] = 0ull;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st_relaxed(&a.ctr->qlen[(ph + 2) & 3], 0u);
        __hip_atomic_store(&a.ctr->heavy_alloc[(ph + 1) & 1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&a.ctr->bu_tiles[(ph + 1) & 1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&a.ctr->bu_tiles_n[(ph + 1) & 1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (n_f == 0) return;  // no vertex carries a bit: nothing can be found
    if (threadIdx.x == 0) s.q.count = 0;
    gather_init(s.g);
    __syncthreads();
    const GlobalQueue out = ms_out_queue(a, ph);
    const unsigned long long* cur = a.front[ph % 3];

    // Top-down push over the frontier queue, for the bits in top-down mode.
    const int32_t* qin = a.queue[ph % 3];
    for (uint32_t base = blockIdx.x * BFS_THREADS; base < n_f; base += gridDim.x * BFS_THREADS) {
        const uint32_t i = base + threadIdx.x;
        int32_t u = -1, start = 0, deg = 0;
        if (i < n_f) {
            const int32_t x = qin[i];
            if ((uint32_t)x < (uint32_t)a.n) {
                if (cur[x] & ~a.bu[a.gid[x]]) {
                    u = x;
                    start = a.offs[x];
                    deg = a.offs[x + 1] - start;
                }
            } else {
                st_relaxed(&a.ctr->corrupt, 1u);
            }
        }
        ms_defer_heavy(a, ph, u, deg);
        auto sink = [&](bool valid, int32_t v, int32_t src) { ms_push_edge(a, ph, s, out, valid, v, src); };
        tiered_gather(a.cols, u, start, deg, a.cta_th, a.wave_th, s.g, sink);
        lq_flush<false>(s.q, out, LQ_CAP / 2);
    }
    lq_flush<false>(s.q, out, 1);
}

// Bottom-up pull of level ph for the bits in bottom-up mode: each vertex of
// such a graph scans its neighbours until every wanted bit is found.  Bits
// are disjoint from the push's, but both OR into visited[v] and next[v], so
// both use atomics there.  A neighbour's 64-bit mask is loaded only when the
// frontier bitmap (one bit per vertex: "carries any bit this level", built
// by k_ms_count, small enough to stay in L2) says it is in some frontier.
//
// Each wave works alone on tiles of MS_TILE = 64 consecutive vertices of one
// graph (lane l = vertex base + l), loading the next tile's visited masks
// and row offsets one tile ahead, and skipping a tile with nothing wanted:
//  * stage 1: each lane checks the first MS_FIRST neighbours of its vertex
//    (loads issued together);
//  * stage 2: vertices still missing wanted bits are scanned by MS_GROUP-lane
//    groups, MS_STRIDE neighbours per lane per step (loads issued together).
// Each new bit is credited to the first neighbour (in list order) that
// carries it, as a credit event (neighbour, bits): stage 1's stay in the
// vertex's lane, stage 2's go to LDS (MS_EVENTS per vertex; further ones
// write their labels at once).  The tile's labels are then written one query
// (bit) at a time, lane l storing vertex base + l, so a wave writes 256
// contiguous bytes per label array instead of scattering one word per query
// per vertex over up to 64 label arrays.
constexpr int MS_FIRST = 4;
constexpr int MS_GROUP = 8;
constexpr int MS_GROUPS = BFS_THREADS / MS_GROUP;
constexpr int MS_STRIDE = 4;
constexpr int MS_EVENTS = 4;
constexpr int MS_QBUF = 128;  // per-wave buffer of discovered vertices
static_assert(MS_TILE == kWave, "a pull tile is one vertex per lane");

struct MsPullWave {
    int32_t vtx[kWave];    // pending vertex (by slot)
    int32_t start[kWave];  // its next edge to scan
    int32_t end[kWave];
    unsigned long long want[kWave];
    unsigned long long got[kWave];
    uint32_t nev[kWave];
    int32_t ev_u[MS_EVENTS][kWave];
    unsigned long long ev_b[MS_EVENTS][kWave];
    int32_t qbuf[MS_QBUF];
};
struct MsPullShared {
    int2 tab[MS_TILE_TABLE];  // (graph, first tile) of the graphs with bottom-up slots
    MsPullWave w[BFS_WAVES];
    uint32_t qn[BFS_WAVES];
    uint32_t qbase[BFS_WAVES];
};

__device__ __forceinline__ bool ms_bm_test(const uint32_t* bm, int32_t u) { return (bm[u >> 5] >> (u & 31)) & 1u; }

// LDS written by some lanes of a wave and read by others after it: keep the
// compiler from reordering across this point (a wave's LDS operations
// complete in program order).
__device__ __forceinline__ void ms_wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// OR over the W lanes of this lane's aligned group, and over the group's
// lanes below this one.
template <int W>
__device__ __forceinline__ unsigned long long group_or(unsigned long long x) {
#pragma unroll
    for (int d = W / 2; d >= 1; d >>= 1) x |= __shfl_xor(x, d, kWave);
    return x;
}
template <int W>
__device__ __forceinline__ unsigned long long group_excl_or(unsigned long long x) {
    const int sub = lane_id() & (W - 1);
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        const unsigned long long y = __shfl_up(x, d, kWave);
        if (sub >= d) x |= y;
    }
    const unsigned long long e = __shfl_up(x, 1, kWave);
    return sub == 0 ? 0ull : e;
}

// Tile t of the level: its first vertex, end and graph (-1: look up per
// vertex).  Tiles never straddle graphs in table mode.
__device__ __forceinline__ void ms_pull_tile(const MsArgs& a, const MsPullShared& s, bool table, int32_t ne,
                                             uint32_t t, int32_t& base, int32_t& end, int32_t& tg) {
    tg = -1;
    if (table) {
        int lo = 0, hi = ne - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((uint32_t)s.tab[mid].y <= t) lo = mid;
            else hi = mid - 1;
        }
        tg = s.tab[lo].x;
        if ((uint32_t)tg < (uint32_t)a.G) {
            base = a.vbase[tg] + (int32_t)(t - (uint32_t)s.tab[lo].y) * MS_TILE;
            end = min(base + MS_TILE, a.vbase[tg + 1]);
        } else {
            if (lane_id() == 0) st_relaxed(&a.ctr->corrupt, 1u);
            tg = -1;
            base = end = 0;
        }
    } else {
        base = (int32_t)t * MS_TILE;
        end = min(base + MS_TILE, a.n);
    }
}

__device__ void ms_pull_body(const MsArgs& a, int ph, MsPullShared& s) {
    uint32_t n_f = ld_relaxed(&a.ctr->qlen[ph & 3]);
    if (n_f > a.qcap) n_f = 0;
    if (n_f == 0 || !ld_relaxed(&a.ctr->any_bu[ph & 1])) return;
    // Tiles: with the level's wide-form graph table (k_ms_direction lists
    // every graph with more than MS_WIDE_BITS bottom-up slots, in tile
    // order), tile t lies in graph tab[e].x for the last e with
    // tab[e].y <= t.  A table too long for LDS: tile t is vertices
    // [t, t+1) * MS_TILE of the pack, with every graph's bottom-up bits; a
    // graph of the narrow form is then pulled twice, the second pass finding
    // only bits the first could not (correct, just slower).
    const unsigned long long tl =
        __hip_atomic_load(&a.ctr->bu_tiles[ph & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int32_t ne = (int32_t)(tl >> 32);
    const bool table = ne <= MS_TILE_TABLE && ne <= a.G;
    if (ne == 0) return;  // no graph pulls in the wide form
    for (int i = threadIdx.x; table && i < ne; i += BFS_THREADS)
        s.tab[i] = a.bu_list[(size_t)(2 + (ph & 1)) * a.G + i];
    const uint32_t ntiles = table ? (uint32_t)tl : (uint32_t)(((int64_t)a.n + MS_TILE - 1) / MS_TILE);
    __syncthreads();
    const unsigned long long* cur = a.front[ph % 3];
    unsigned long long* next = a.front[(ph + 1) % 3];
    const uint32_t* fbm = a.fbm[ph & 1];
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    const int grp = lane / MS_GROUP, sub = lane % MS_GROUP;
    MsPullWave& W = s.w[wid];
    const GlobalQueue out = ms_out_queue(a, ph);
    uint32_t wq = 0;  // vertices in this wave's buffer (wave-uniform)
    int32_t cg = -1;   // graph whose slot labels the lanes hold (slb, sdv)
    long long slb = 0;
    int32_t sdv = -1;

    const uint32_t tstep = gridDim.x * BFS_WAVES;
    uint32_t t = blockIdx.x * BFS_WAVES + wid;
    // The next tile's masks and row offsets are loaded one tile ahead.
    int32_t nbase = 0, nend = 0, ntg = -1;
    unsigned long long nm = 0, nvis = 0;
    int32_t noff0 = 0, noff1 = 0;
    auto prefetch = [&](uint32_t tt) {
        nm = 0;
        if (tt >= ntiles) return;
        ms_pull_tile(a, s, table, ne, tt, nbase, nend, ntg);
        const int32_t v = nbase + lane;
        if (v < nend) {
            nm = a.bu[ntg >= 0 ? ntg : a.gid[v]];
            nvis = a.visited[v];
            noff0 = a.offs[v];
            noff1 = a.offs[v + 1];
        }
    };
    prefetch(t);
    for (; t < ntiles; t += tstep) {
        const int32_t base = nbase, end = nend, tg = ntg;
        const unsigned long long want = nm & ~nvis;
        const int32_t start = noff0, deg = want ? noff1 - noff0 : 0;
        prefetch(t + tstep);
        if (__ballot(want != 0ull) == 0ull) continue;
        const int32_t v = base + lane;

        // Stage 1.
        unsigned long long got = 0;
        int32_t e1u[MS_FIRST];
        unsigned long long e1b[MS_FIRST];
        {
            int32_t u[MS_FIRST];
            unsigned long long f[MS_FIRST];
#pragma unroll
            for (int k = 0; k < MS_FIRST; ++k) u[k] = (k < deg) ? a.cols[start + k] : -1;
#pragma unroll
            for (int k = 0; k < MS_FIRST; ++k) f[k] = (u[k] >= 0 && ms_bm_test(fbm, u[k])) ? cur[u[k]] & want : 0ull;
#pragma unroll
            for (int k = 0; k < MS_FIRST; ++k) {
                e1u[k] = u[k];
                e1b[k] = f[k] & ~got;
                got |= f[k];
            }
        }

        // Stage 2: the wave's unresolved lists, one per MS_GROUP-lane group.
        const bool pend = want != 0ull && got != want && deg > MS_FIRST;
        const unsigned long long pm = __ballot(pend);
        int slot = 0;
        int nev2 = 0;
        int32_t e2u[MS_EVENTS];
        unsigned long long e2b[MS_EVENTS];
#pragma unroll
        for (int k = 0; k < MS_EVENTS; ++k) {
            e2u[k] = -1;
            e2b[k] = 0ull;
        }
        if (pm) {
            slot = lane_prefix(pm);
            if (pend) {
                W.vtx[slot] = v;
                W.start[slot] = start + MS_FIRST;
                W.end[slot] = start + deg;
                W.want[slot] = want;
                W.got[slot] = got;
                W.nev[slot] = 0u;
            }
            ms_wave_lds_sync();
            const int np = __popcll(pm);
            for (int i = grp; i < np; i += kWave / MS_GROUP) {
                const int32_t e1 = W.end[i];
                const unsigned long long w = W.want[i];
                unsigned long long g = W.got[i];
                for (int32_t e0 = W.start[i]; e0 < e1 && g != w; e0 += MS_GROUP * MS_STRIDE) {
                    int32_t uu[MS_STRIDE];
                    unsigned long long ff[MS_STRIDE];
#pragma unroll
                    for (int k = 0; k < MS_STRIDE; ++k) {
                        const int32_t e = e0 + k * MS_GROUP + sub;
                        uu[k] = e < e1 ? a.cols[e] : -1;
                    }
#pragma unroll
                    for (int k = 0; k < MS_STRIDE; ++k)
                        ff[k] = (uu[k] >= 0 && ms_bm_test(fbm, uu[k])) ? cur[uu[k]] & w : 0ull;
#pragma unroll
                    for (int k = 0; k < MS_STRIDE; ++k) {
                        const unsigned long long f = ff[k] & ~g;
                        const unsigned long long mine = f & ~group_excl_or<MS_GROUP>(f);
                        if (mine) {
                            const uint32_t n = atomicAdd(&W.nev[i], 1u);
                            if (n < (uint32_t)MS_EVENTS) {
                                W.ev_u[n][i] = uu[k];
                                W.ev_b[n][i] = mine;
                            } else {
                                // Out of event slots: this credit's labels now.
                                ms_record(a, ph, W.vtx[i], uu[k], mine);
                            }
                        }
                        g |= group_or<MS_GROUP>(f);
                    }
                }
                if (sub == 0) W.got[i] = g;
            }
            ms_wave_lds_sync();
            if (pend) {
                got = W.got[slot];
                nev2 = (int)min(W.nev[slot], (uint32_t)MS_EVENTS);
#pragma unroll
                for (int k = 0; k < MS_EVENTS; ++k) {
                    if (k < nev2) {
                        e2u[k] = W.ev_u[k][slot];
                        e2b[k] = W.ev_b[k][slot];
                    }
                }
            }
            ms_wave_lds_sync();  // the next tile rewrites the scratch
        }

        // Claim the new bits and buffer the vertices that join the next
        // frontier.
        bool enq = false;
        if (got) {
            atomic_or_u64(&a.visited[v], got);
            enq = atomic_or_u64(&next[v], got) == 0ull;
        }
        const unsigned long long em = __ballot(enq);
        if (em) {
            const uint32_t nf = (uint32_t)__popcll(em);
            if (wq + nf > (uint32_t)MS_QBUF) {  // buffer full: flush this wave's part now
                uint32_t qb0 = 0;
                if (lane == 0) qb0 = atomic_add_u32(out.len, wq);
                qb0 = (uint32_t)wave_bcast((int)qb0, 0);
                for (uint32_t i = lane; i < wq; i += kWave) {
                    if (qb0 + i < out.cap) out.keys[qb0 + i] = W.qbuf[i];
                    else st_relaxed(out.overflow, 1u);
                }
                wq = 0;
                ms_wave_lds_sync();
            }
            if (enq) W.qbuf[wq + (uint32_t)lane_prefix(em)] = v;
            wq += nf;
        }

        // Labels, one query (bit) at a time.
        unsigned long long staged = 0;
#pragma unroll
        for (int k = 0; k < MS_FIRST; ++k) staged |= e1b[k];
#pragma unroll
        for (int k = 0; k < MS_EVENTS; ++k) staged |= e2b[k];
        // Lane b holds bit b's label offset and depth value for the tile's
        // graph (reloaded when the graph changes), read back per bit with a
        // lane read instead of two dependent loads.
        if (tg >= 0 && tg != cg) {
            cg = tg;
            const int32_t q = a.bitq[tg * 64 + lane];
            slb = q >= 0 ? a.lbase[q] - a.vbase[tg] : 0ll;  // may be negative
            sdv = q >= 0 ? ph + 1 - a.qstart[q] : -1;      // -1: no query in slot
        }
        unsigned long long bits = group_or<kWave>(staged);
        while (bits) {
            const int b = __builtin_amdgcn_readfirstlane(__ffsll(bits) - 1);
            bits &= bits - 1;
            long long lb;
            int32_t dv, gv;
            if (tg >= 0) {
                lb = (long long)(((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(slb >> 32), b) << 32) |
                                 (uint32_t)__builtin_amdgcn_readlane((int)slb, b));
                dv = __builtin_amdgcn_readlane(sdv, b);
                gv = tg;
            } else {
                gv = v < end ? a.gid[v] : 0;
                const int32_t q = a.bitq[gv * 64 + b];
                lb = q >= 0 ? a.lbase[q] - a.vbase[gv] : 0ll;
                dv = q >= 0 ? ph + 1 - a.qstart[q] : -1;
            }
            if (!((staged >> b) & 1ull) || dv < 0) continue;
            int32_t pu = -1;
#pragma unroll
            for (int k = 0; k < MS_FIRST; ++k)
                if ((e1b[k] >> b) & 1ull) pu = e1u[k];
#pragma unroll
            for (int k = 0; k < MS_EVENTS; ++k)
                if ((e2b[k] >> b) & 1ull) pu = e2u[k];
            a.depth_out[lb + v] = dv;
            a.pred_out[lb + v] = pu - a.vbase[gv];
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
        const uint32_t qb0 = total ? atomic_add_u32(out.len, total) : 0u;
        for (int w = 0; w < BFS_WAVES; ++w) s.qbase[w] += qb0;
    }
    __syncthreads();
    for (uint32_t i = lane; i < wq; i += kWave) {
        const uint32_t pos = s.qbase[wid] + i;
        if (pos < out.cap) out.keys[pos] = W.qbuf[i];
        else st_relaxed(out.overflow, 1u);
    }
}

// Narrow form of the pull, for graphs with at most MS_WIDE_BITS slots in
// bottom-up mode: workgroups take tiles of MS_TILE_N consecutive vertices
// (one per thread); stage 1 checks each vertex's first MS_FIRST neighbours,
// stage 2 scans the unresolved lists with MS_GROUP-lane groups, one
// neighbour per lane per step, and every credit's labels are written as it
// is found (few bits per vertex: the stores stay few).  At 8 waves per SIMD
// it beats the wide form when most vertices want only a few bits.
struct MsPullNarrowShared {
    LdsKeyQueue q;
    int2 tab[MS_TILE_TABLE];  // (graph, first tile) of the graphs with bottom-up slots
    int32_t pv[BFS_THREADS];  // unresolved vertex (tile-local index)
    int32_t pe[BFS_THREADS];  // next edge to scan
    int32_t pend_end[BFS_THREADS];
    unsigned long long pwant[BFS_THREADS];
    unsigned long long got[BFS_THREADS];
    uint32_t npend;
};

__device__ void ms_pull_narrow_body(const MsArgs& a, int ph, MsPullNarrowShared& s) {
    uint32_t n_f = ld_relaxed(&a.ctr->qlen[ph & 3]);
    if (n_f > a.qcap) n_f = 0;
    if (n_f == 0 || !ld_relaxed(&a.ctr->any_bu[ph & 1])) return;
    // Tiles: with the level's narrow-form graph table (graphs with at most
    // MS_WIDE_BITS bottom-up slots, in tile order), tile t lies in graph
    // tab[e].x for the last e with tab[e].y <= t.  A table too long for LDS:
    // tile t is vertices [t, t+1) * MS_TILE_N of the pack, with every graph's
    // bottom-up bits (the wide pull that follows then finds only what is
    // left of its graphs).
    const unsigned long long tl =
        __hip_atomic_load(&a.ctr->bu_tiles_n[ph & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int32_t ne = (int32_t)(tl >> 32);
    const bool table = ne <= MS_TILE_TABLE && ne <= a.G;
    for (int i = threadIdx.x; table && i < ne; i += BFS_THREADS) s.tab[i] = a.bu_list[(size_t)(ph & 1) * a.G + i];
    const uint32_t ntiles = table ? (uint32_t)tl : (uint32_t)(((int64_t)a.n + MS_TILE_N - 1) / MS_TILE_N);
    if (threadIdx.x == 0) s.q.count = 0;
    __syncthreads();
    const GlobalQueue out = ms_out_queue(a, ph);
    const unsigned long long* cur = a.front[ph % 3];
    unsigned long long* next = a.front[(ph + 1) % 3];
    const uint32_t* fbm = a.fbm[ph & 1];
    const int tid = threadIdx.x;
    const int grp = tid / MS_GROUP, sub = tid % MS_GROUP;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int32_t base, end, tg = -1;
        if (table) {
            int lo = 0, hi = ne - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((uint32_t)s.tab[mid].y <= t) lo = mid;
                else hi = mid - 1;
            }
            tg = s.tab[lo].x;
            if ((uint32_t)tg < (uint32_t)a.G) {
                base = a.vbase[tg] + (int32_t)(t - (uint32_t)s.tab[lo].y) * MS_TILE_N;
                end = min(base + MS_TILE_N, a.vbase[tg + 1]);
            } else {
                if (tid == 0) st_relaxed(&a.ctr->corrupt, 1u);
                tg = -1;
                base = end = 0;
            }
        } else {
            base = (int32_t)t * MS_TILE_N;
            end = min(base + MS_TILE_N, a.n);
        }
        if (tid == 0) s.npend = 0;
        const int32_t v = base + tid;
        unsigned long long want = 0, got = 0;
        int32_t start = 0, deg = 0;
        if (v < end) {
            const unsigned long long m = a.bu[tg >= 0 ? tg : a.gid[v]];
            if (m) {
                want = m & ~a.visited[v];
                if (want) {
                    start = a.offs[v];
                    deg = a.offs[v + 1] - start;
                }
            }
        }
        __syncthreads();  // npend reset visible; the previous tile is done with the LDS arrays
        // Stage 1: the first MS_FIRST neighbours, loads issued together.
        if (want != 0ull && deg > 0) {
            int32_t u[MS_FIRST];
#pragma unroll
            for (int k = 0; k < MS_FIRST; ++k) u[k] = (k < deg) ? a.cols[start + k] : -1;
#pragma unroll
            for (int k = 0; k < MS_FIRST; ++k) {
                if (u[k] >= 0 && ms_bm_test(fbm, u[k])) {
                    const unsigned long long f = cur[u[k]] & want & ~got;
                    if (f) {
                        got |= f;
                        ms_record(a, ph, v, u[k], f);
                    }
                }
            }
            if (got != want && deg > MS_FIRST) {
                const uint32_t slot = atomicAdd(&s.npend, 1u);
                s.pv[slot] = tid;
                s.pe[slot] = start + MS_FIRST;
                s.pend_end[slot] = start + deg;
                s.pwant[slot] = want;
            }
        }
        s.got[tid] = got;
        __syncthreads();
        // Stage 2: groups of MS_GROUP lanes scan the remaining lists.
        const uint32_t np = s.npend;
        for (uint32_t i = grp; i < np; i += MS_GROUPS) {
            const int32_t lt = s.pv[i];
            const int32_t vv = base + lt;
            const int32_t e1 = s.pend_end[i];
            const unsigned long long w = s.pwant[i];
            unsigned long long g = s.got[lt];
            for (int32_t e0 = s.pe[i]; e0 < e1 && g != w; e0 += MS_GROUP) {
                const int32_t e = e0 + sub;
                int32_t u = -1;
                unsigned long long f = 0;
                if (e < e1) {
                    u = a.cols[e];
                    if (ms_bm_test(fbm, u)) f = cur[u] & w & ~g;
                }
                const unsigned long long mine = f & ~group_excl_or<MS_GROUP>(f);
                if (mine) ms_record(a, ph, vv, u, mine);
                g |= group_or<MS_GROUP>(f);
            }
            if (sub == 0) s.got[lt] = g;
        }
        __syncthreads();
        got = s.got[tid];
        bool enq = false;
        if (got) {
            atomic_or_u64(&a.visited[v], got);
            enq = atomic_or_u64(&next[v], got) == 0ull;
        }
        lq_append<false>(s.q, out, enq, v, 0);
        lq_flush<false>(s.q, out, LQ_CAP / 2);
    }
    lq_flush<false>(s.q, out, 1);
}

// Grid tier of level ph's push: workgroups take chunks grid-stride, one edge
// per thread.  Each chunk is checked against the entry it names, so a stale
// map word never indexes outside a list.
__device__ void ms_heavy_body(const MsArgs& a, int ph, MsShared& s) {
    const unsigned long long alloc =
        __hip_atomic_load(&a.ctr->heavy_alloc[ph & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t H = min((uint32_t)(alloc >> 32), a.heavy_cap);
    const uint32_t C = min((uint32_t)alloc, a.chunk_cap);
    if (H == 0 || C == 0) return;
    if (threadIdx.x == 0) s.q.count = 0;
    __syncthreads();
    const GlobalQueue out = ms_out_queue(a, ph);
    for (uint32_t c = blockIdx.x; c < C; c += gridDim.x) {
        const int32_t entry = a.chunk_map[c];
        int32_t u = -1, e = 0;
        bool valid = false;
        if ((uint32_t)entry < H) {
            const int2 eu = a.heavy[entry];
            if ((uint32_t)eu.x < (uint32_t)a.n && (uint32_t)eu.y <= c) {
                const int32_t s0 = a.offs[eu.x], deg = a.offs[eu.x + 1] - s0;
                const long long off = (long long)(c - (uint32_t)eu.y) * HEAVY_CHUNK + threadIdx.x;
                if (off < deg) {
                    u = eu.x;
                    e = s0 + (int32_t)off;
                    valid = true;
                }
            } else {
                st_relaxed(&a.ctr->corrupt, 1u);
            }
        }
        const int32_t v = valid ? a.cols[e] : -1;
        ms_push_edge(a, ph, s, out, valid, v, u);
        lq_flush<false>(s.q, out, LQ_CAP / 2);
    }
    lq_flush<false>(s.q, out, 1);
}

// Per (graph, slot) count and degree sum of the vertices discovered by level
// ph (the next frontier queue; next[v] holds exactly v's new bits).  Each
// wave takes one contiguous stretch of the queue (entries appended together
// mostly share a graph) and, per chunk of 64 entries and per graph present in
// it, transposes the lanes' bit masks so that lane b holds the set of lanes
// carrying bit b.  Running totals are kept while the graph stays the same.
__device__ void ms_count_body(const MsArgs& a, int ph) {
    uint32_t n_q = ld_relaxed(&a.ctr->qlen[(ph + 1) & 3]);
    // Frontier bitmap of level ph + 1: set entry by entry from the queue while
    // it is short (the array is all zero: its bits were those of level ph - 1,
    // cleared below one level ago), else rewritten word by word from the
    // masks after the counts.
    const bool dense = n_q > a.qcap || (int64_t)n_q * MS_SPARSE_DIV >= (int64_t)a.n;
    if (n_q > a.qcap) n_q = 0;
    const int32_t* q = a.queue[(ph + 1) % 3];
    const unsigned long long* nxt = a.front[(ph + 1) % 3];
    uint32_t* fbm_next = a.fbm[(ph + 1) & 1];
    const int lane = lane_id();
    const int64_t wave = (int64_t)blockIdx.x * BFS_WAVES + threadIdx.x / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * BFS_WAVES;
    const int64_t per = (((int64_t)n_q + nwaves - 1) / nwaves + kWave - 1) / kWave * kWave;
    const int64_t lo = wave * per, hi = min((int64_t)n_q, lo + per);
    int32_t accg = -1;
    uint32_t accn = 0;
    unsigned long long accm = 0;
    for (int64_t base = lo; base < hi; base += kWave) {
        const int64_t i = base + lane;
        int32_t g = -1, deg = 0;
        unsigned long long bits = 0;
        if (i < hi) {
            const int32_t v = q[i];
            if ((uint32_t)v < (uint32_t)a.n) {
                g = a.gid[v];
                bits = nxt[v];
                deg = a.offs[v + 1] - a.offs[v];
                if (!dense) atomicOr(&fbm_next[v >> 5], 1u << (v & 31));
            } else {
                st_relaxed(&a.ctr->corrupt, 1u);
            }
        }
        bool todo = g >= 0;
        for (uint64_t rest = __ballot(todo); rest; rest = __ballot(todo)) {
            const int32_t g1 = __shfl(g, __ffsll((unsigned long long)rest) - 1, kWave);
            const bool mine = todo && g == g1;
            todo = todo && !mine;
            const unsigned long long y = wave_transpose64(mine ? bits : 0ull);
            unsigned long long m = 0;
#pragma unroll 8
            for (int l = 0; l < kWave; ++l) {
                const int d = __shfl(deg, l, kWave);
                if ((y >> l) & 1ull) m += (unsigned long long)d;
            }
            if (g1 != accg) {
                if (accg >= 0 && accn) {
                    atomic_add_u32(&a.cnt_n[(size_t)accg * 64 + lane], accn);
                    atomic_add_u64(&a.cnt_m[(size_t)accg * 64 + lane], accm);
                }
                accg = g1;
                accn = 0;
                accm = 0;
            }
            accn += (uint32_t)__popcll(y);
            accm += m;
        }
    }
    if (accg >= 0 && accn) {
        atomic_add_u32(&a.cnt_n[(size_t)accg * 64 + lane], accn);
        atomic_add_u64(&a.cnt_m[(size_t)accg * 64 + lane], accm);
    }
    const int64_t words = ((int64_t)a.n + 31) / 32;
    if (dense) {
        // bit v = (next[v] != 0); each wave writes the two words of its 64
        // vertices.
        for (int64_t vb = wave * kWave; vb < a.n; vb += nwaves * kWave) {
            const int64_t v = vb + lane;
            const uint64_t m = __ballot(v < a.n && nxt[v] != 0ull);
            if ((lane & 31) == 0) {
                const int64_t w = v >> 5;
                if (w < words) fbm_next[w] = (lane == 0) ? (uint32_t)m : (uint32_t)(m >> 32);
            }
        }
    }
    // Level ph's bitmap is no longer read (the pull is done); zero it for
    // level ph + 2.
    uint32_t* fbm_cur = a.fbm[ph & 1];
    for (int64_t w = (int64_t)blockIdx.x * BFS_THREADS + threadIdx.x; w < words; w += (int64_t)gridDim.x * BFS_THREADS)
        fbm_cur[w] = 0u;
}

// One wave per graph, lane b = slot b.  The slot's level counts decide its
// direction for level ph + 1: top-down -> bottom-up when the new frontier's
// degree sum m_f exceeds m_u / alpha (m_u = edges of the graph not yet
// reached by this query), bottom-up -> top-down when the frontier has shrunk
// below V_g / beta.  Consumes (zeroes) the counts.
//
// Retirement happens here too: a slot in use whose level discovered nothing
// has finished.  Its active bit is cleared (later levels, also those of the
// same captured replay, skip it) and the query's level count is written to
// the pinned host word hdone[q]; the host frees the slot and its labels at
// its next level boundary (k_ms_release clears the slot's visited bits).
// One graph per wave, waves grid-strided over the graphs.
__device__ void ms_direction_body(const MsArgs& a, int ph) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed(&a.ctr->any_bu[ph & 1], 0u);
    const int b = lane_id();
    for (int32_t g = (int32_t)(blockIdx.x * BFS_WAVES + threadIdx.x / kWave); g < a.G;
         g += (int32_t)(gridDim.x * BFS_WAVES)) {
        const size_t k = (size_t)g * 64 + b;
        const uint32_t n = a.cnt_n[k];
        const unsigned long long m = a.cnt_m[k];
        a.cnt_n[k] = 0u;
        a.cnt_m[k] = 0ull;
        const bool occupied = (a.active[g] >> b) & 1ull;
        const bool was_bu = (a.bu[g] >> b) & 1ull;
        const long long Vg = a.vbase[g + 1] - a.vbase[g];
        const long long Eg = a.ebase[g + 1] - a.ebase[g];
        bool to_bu = false;
        if (occupied && n) {
            const unsigned long long seen = a.m_seen[k] + m;
            a.m_seen[k] = seen;
            if (!was_bu) {
                long long m_u = Eg - (long long)seen;
                if (m_u < 0) m_u = 0;
                to_bu = (long long)m * (long long)a.alpha > m_u;
            } else {
                to_bu = !((long long)n * (long long)a.beta < Vg && n < a.nf_prev[k]);
            }
        }
        a.nf_prev[k] = n;
        const unsigned long long bum = __ballot(to_bu);
        const bool finished = occupied && n == 0u;
        const unsigned long long done = __ballot(finished);
        if (finished) {
            const int32_t q = a.bitq[k];
            if (q >= 0) a.hdone[q] = ph + 1 - a.qstart[q];
        }
        if (b == 0) {
            a.bu[g] = bum;
            if (done) atomic_and_u64(&a.active[g], ~done);
            if (bum) {
                // List the graph for the next level's pull, in the wide or
                // the narrow table by its bottom-up slot count: one 64-bit
                // atomic gives both its table entry and its first tile, so
                // each table is in tile order.
                st_relaxed(&a.ctr->any_bu[(ph + 1) & 1], 1u);
                const int cls = __popcll(bum) > MS_WIDE_BITS ? 1 : 0;
                const int tile = cls ? MS_TILE : MS_TILE_N;
                const uint32_t nt = (uint32_t)((Vg + tile - 1) / tile);
                const unsigned long long t = __hip_atomic_fetch_add(
                    cls ? &a.ctr->bu_tiles[(ph + 1) & 1] : &a.ctr->bu_tiles_n[(ph + 1) & 1],
                    (1ull << 32) | (unsigned long long)nt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t e = (uint32_t)(t >> 32);
                if (e < (uint32_t)a.G)
                    a.bu_list[(size_t)(cls * 2 + ((ph + 1) & 1)) * a.G + e] = make_int2(g, (int)(uint32_t)t);
            }
        }
    }
    // The level's last step: publish the error flags beside the finished
    // queries' words (the host reads both after one stream synchronisation,
    // no copies).
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.hctr->overflow = ld_relaxed(&a.ctr->overflow);
        a.hctr->corrupt = ld_relaxed(&a.ctr->corrupt);
        a.hctr->bar_abort = ld_relaxed(&a.ctr->bar_abort);
    }
}

// The level kernels: each wraps its body (launched one after another, or
// replayed from a captured graph with the phase read from device memory).
__global__ __launch_bounds__(BFS_THREADS) void k_ms_push(MsArgs a, int ph) {
    __shared__ MsShared s;
    ms_push_body(a, ms_phase(a, ph), s);
}
__global__ __launch_bounds__(BFS_THREADS) void k_ms_heavy(MsArgs a, int ph) {
    __shared__ MsShared s;
    ms_heavy_body(a, ms_phase(a, ph), s);
}
__global__ __launch_bounds__(BFS_THREADS) void k_ms_pull(MsArgs a, int ph) {
    __shared__ MsPullShared s;
    ms_pull_body(a, ms_phase(a, ph), s);
}
__global__ __launch_bounds__(BFS_THREADS) void k_ms_pull_narrow(MsArgs a, int ph) {
    __shared__ MsPullNarrowShared s;
    ms_pull_narrow_body(a, ms_phase(a, ph), s);
}
__global__ __launch_bounds__(BFS_THREADS) void k_ms_count(MsArgs a, int ph) { ms_count_body(a, ms_phase(a, ph)); }
__global__ __launch_bounds__(BFS_THREADS) void k_ms_direction(MsArgs a, int ph) {
    ms_direction_body(a, ms_phase(a, ph));
}

// Persistent form: `levels` whole levels in one launch of a co-resident grid,
// the six steps of a level separated by grid barriers (fenced mode: the
// steps pass masks, queues and counters through plain accesses).  Queries
// that finish retire on the device as in the other forms; a level after the
// last query finished costs only its barriers.
union MsFusedShared {
    MsShared push;
    MsPullShared pull;
    MsPullNarrowShared pull_n;
};
__global__ __launch_bounds__(BFS_THREADS) void k_ms_fused(MsArgs a, int ph0, int levels) {
    __shared__ MsFusedShared s;
    __shared__ uint32_t s_word;
    const GridBarrier bar{a.bar_arrive, &a.ctr->bar_release, &a.ctr->bar_abort, 0u};
    uint32_t epoch = barrier_base_epoch(bar);
    for (int l = 0; l < levels; ++l) {
        const int ph = ph0 + l;
        ms_push_body(a, ph, s.push);
        if (!grid_sync(bar, ++epoch, &s_word)) break;
        ms_heavy_body(a, ph, s.push);
        if (!grid_sync(bar, ++epoch, &s_word)) break;
        ms_pull_narrow_body(a, ph, s.pull_n);
        if (!grid_sync(bar, ++epoch, &s_word)) break;
        ms_pull_body(a, ph, s.pull);
        if (!grid_sync(bar, ++epoch, &s_word)) break;
        ms_count_body(a, ph);
        if (!grid_sync(bar, ++epoch, &s_word)) break;
        ms_direction_body(a, ph);
        if (!grid_sync(bar, ++epoch, &s_word)) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.hctr->bar_abort = ld_relaxed(&a.ctr->bar_abort);
}

// Last node of a captured level loop: moves the replay's first phase on by
// the `levels` levels the replay ran.
__global__ void k_ms_advance(MsArgs a, int levels) {
    if (threadIdx.x == 0) st_relaxed((uint32_t*)a.dph, ld_relaxed((const uint32_t*)a.dph) + (uint32_t)levels);
}

// Sets the device phase word outside a captured loop (when launches and
// replays are mixed on one engine).
__global__ void k_ms_set_phase(MsArgs a, int ph) {
    if (threadIdx.x == 0) st_relaxed((uint32_t*)a.dph, (uint32_t)ph);
}


// Per admitted query: graph slot, label base and start level.
struct MsSeed {
    int32_t q;
    int32_t src;  // global id
    int32_t bit;
    int32_t qstart;
    long long lbase;
};

// Admits the queries of one level boundary, one row of workgroups per query
// (blockIdx.y): the row initialises the query's label slot (depth and pred
// -1, the source 0 and -2), and its first thread registers the query (slot
// table, active mask, label base, start level, top-down mode, fresh
// direction statistics) and seeds level ph: source bit in visited and the
// current mask array, and the source appended to the current frontier queue
// if it was not already on it.  The slot table and masks are only written
// here, in k_ms_release and in k_ms_direction, never copied from the host.
__global__ __launch_bounds__(256) void k_ms_admit(MsArgs a, int ph, const MsSeed* seeds, int count) {
    const int i = blockIdx.y;
    if (i >= count) return;
    const MsSeed sd = seeds[i];
    const int32_t q = sd.q, v = sd.src;
    if ((uint32_t)v >= (uint32_t)a.n || (uint32_t)sd.bit >= 64u) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed(&a.ctr->corrupt, 1u);
        return;
    }
    const int32_t g = a.gid[v];
    const int32_t vb = a.vbase[g], vg = a.vbase[g + 1] - vb, ls = v - vb;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < vg; k += gridDim.x * blockDim.x) {
        a.depth_out[sd.lbase + k] = (k == ls) ? 0 : -1;
        a.pred_out[sd.lbase + k] = (k == ls) ? -2 : -1;
    }
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned long long bit = 1ull << sd.bit;
    const size_t k = (size_t)g * 64 + sd.bit;
    a.bitq[k] = q;
    a.lbase[q] = sd.lbase;
    a.qstart[q] = sd.qstart;
    a.nf_prev[k] = 1u;
    a.m_seen[k] = 0ull;
    atomic_and_u64(&a.bu[g], ~bit);
    atomic_or_u64(&a.active[g], bit);
    atomic_or_u64(&a.visited[v], bit);
    const unsigned long long old = atomic_or_u64(&a.front[ph % 3][v], bit);
    if (old == 0ull) {
        const uint32_t pos = atomic_add_u32(&a.ctr->qlen[ph & 3], 1u);
        if (pos < a.qcap)
            a.queue[ph % 3][pos] = v;
        else
            st_relaxed(&a.ctr->overflow, 1u);
    }
}

// Slots of one graph retired at one level boundary.
struct MsRelease {
    int32_t g;
    int32_t pad;
    unsigned long long mask;
};

// Releases retired slots, one row of workgroups per graph (blockIdx.y):
// slot table entries, active and bottom-up bits, and the bits in visited
// over the graph's vertex range.
__global__ __launch_bounds__(256) void k_ms_release(MsArgs a, const MsRelease* rel, int count) {
    const int r = blockIdx.y;
    if (r >= count) return;
    const MsRelease R = rel[r];
    if ((uint32_t)R.g >= (uint32_t)a.G) return;
    const unsigned long long keep = ~R.mask;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (unsigned long long b = R.mask; b; b &= b - 1) a.bitq[(size_t)R.g * 64 + (__ffsll(b) - 1)] = -1;
        atomic_and_u64(&a.active[R.g], keep);
        atomic_and_u64(&a.bu[R.g], keep);
    }
    const int32_t v0 = a.vbase[R.g], v1 = a.vbase[R.g + 1];
    for (int32_t v = v0 + blockIdx.x * blockDim.x + threadIdx.x; v < v1; v += gridDim.x * blockDim.x)
        a.visited[v] &= keep;
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
    // Reference counted: each batched engine built on the pack holds a
    // reference, so destroying the pack before its engines is safe.
    std::atomic<int> refs{1};
    b40c_graph* packed = nullptr;
    int32_t ngraphs = 0;
    std::vector<int32_t> vbase;  // ngraphs + 1
    std::vector<int64_t> ebase;  // ngraphs + 1
    b40c::DeviceBuffer d_vbase, d_ebase, d_gid;
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
    b40c::DeviceBuffer visited, front[3], active, bu, cnt_n, cnt_m, nf_prev, m_seen, bitq, lbase, qstart, depth,
        pred, queue[3], counters, dph;
    b40c::DeviceBuffer seed_buf, rel_buf, heavy, chunk_map, fbm, bu_list;
    b40c::MsSeed* h_seeds = nullptr;  // pinned staging
    b40c::MsRelease* h_rel = nullptr;  // pinned staging
    int h_rel_cap = 0;
    int h_seed_cap = 0;
    uint32_t qcap = 0, heavy_cap = 0, chunk_cap = 0;
    b40c::ArenaAlloc arena;
    std::vector<b40c::MsQuery> queries;
    std::vector<int32_t> free_ids;  // ids whose results were taken, reused by later submissions
    std::deque<int32_t> waiting;
    std::vector<unsigned long long> h_active;  // [G]
    std::vector<int32_t> running;              // query ids in flight
    std::vector<int32_t> completed;            // finished, not yet reported
    int32_t ph = 0;
    int32_t alpha = 14, beta = 24, grid = 1024;
    int32_t cohort_gap = 6;        // levels between admissions while queries run
    int32_t last_admit_ph = -1000;
    b40c::MsCounters* h_ctr = nullptr;      // pinned, written by k_ms_direction (overflow, corrupt)
    b40c::MsCounters* h_ctr_dev = nullptr;  // its device address
    int32_t* h_done = nullptr;      // pinned [qmax]: levels of a query once finished (0 while running)
    int32_t* h_done_dev = nullptr;  // its device address
    int32_t qmax = 0;
    // Captured level loop (levels_per_replay > 0): the six level launches of
    // `levels_per_replay` levels and k_ms_advance, replayed as one graph.
    int32_t levels_per_replay = 0;
    bool persistent = false;             // levels_per_replay levels per step in one k_ms_fused launch
    b40c::DeviceBuffer bar_arrive;       // its grid-barrier arrival words
    bool bar_dirty = false;              // a barrier wait timed out: clear the barrier state first
    hipGraph_t loop_graph = nullptr;
    hipGraphExec_t loop_exec = nullptr;
    int32_t loop_levels = 0, loop_grid = 0, loop_alpha = 0, loop_beta = 0;
    int32_t dev_ph = 0;  // value of the device phase word (dph)
    void release_loop() {
        if (loop_exec) (void)hipGraphExecDestroy(loop_exec);
        if (loop_graph) (void)hipGraphDestroy(loop_graph);
        loop_exec = nullptr;
        loop_graph = nullptr;
    }
};

namespace b40c {
namespace {

MsArgs ms_args(b40c_msbfs* m) {
    MsArgs a{};
    a.offs = m->mg->packed->d_offsets();
    a.cols = m->mg->packed->d_cols();
    a.gid = m->mg->d_gid.as<int32_t>();
    a.vbase = m->mg->d_vbase.as<int32_t>();
    a.ebase = m->mg->d_ebase.as<long long>();
    a.visited = m->visited.as<unsigned long long>();
    for (int i = 0; i < 3; ++i) a.front[i] = m->front[i].as<unsigned long long>();
    a.active = m->active.as<unsigned long long>();
    a.bu = m->bu.as<unsigned long long>();
    a.cnt_n = m->cnt_n.as<uint32_t>();
    a.cnt_m = m->cnt_m.as<unsigned long long>();
    a.nf_prev = m->nf_prev.as<uint32_t>();
    a.m_seen = m->m_seen.as<unsigned long long>();
    a.bitq = m->bitq.as<int32_t>();
    a.lbase = m->lbase.as<long long>();
    a.qstart = m->qstart.as<int32_t>();
    a.depth_out = m->depth.as<int32_t>();
    a.pred_out = m->pred.as<int32_t>();
    for (int i = 0; i < 3; ++i) a.queue[i] = m->queue[i].as<int32_t>();
    a.qcap = m->qcap;
    const size_t words = ((size_t)m->n + 31) / 32;
    a.fbm[0] = m->fbm.as<uint32_t>();
    a.fbm[1] = m->fbm.as<uint32_t>() + words;
    a.bu_list = m->bu_list.as<int2>();
    a.hctr = m->h_ctr_dev;
    a.hdone = m->h_done_dev;
    a.dph = nullptr;
    a.bar_arrive = m->bar_arrive.as<uint32_t>();
    a.heavy = m->heavy.as<int2>();
    a.chunk_map = m->chunk_map.as<int32_t>();
    a.heavy_cap = m->heavy_cap;
    a.chunk_cap = m->chunk_cap;
    a.ctr = m->counters.as<MsCounters>();
    a.n = m->n;
    a.G = m->G;
    a.alpha = m->alpha;
    a.beta = m->beta;
    a.cta_th = 256;
    a.wave_th = 64;
    a.grid_th = 128;
    return a;
}

unsigned grid_for_n(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// Admits waiting queries while a slot of their graph and label memory are free.
b40c_status ms_admit(b40c_msbfs* m) {
    // Cohorts: queries that start together share their expensive first
    // levels (top-down hub expansion, the first bottom-up sweep); starting a
    // few queries at every boundary would pay those levels again each time.
    if (m->waiting.empty()) return B40C_OK;
    if (!m->running.empty() && m->ph - m->last_admit_ph < m->cohort_gap) return B40C_OK;
    std::vector<MsSeed> seeds;
    std::deque<int32_t> still;
    int64_t max_vg = 1;
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
        // The id may have been used before: its done word restarts at 0
        // before the launch that admits it, so the device's later write is
        // the only one that can follow.
        m->h_done[q] = 0;
        m->h_active[g] |= 1ull << b;
        max_vg = std::max(max_vg, vg);
        seeds.push_back(MsSeed{q, m->mg->vbase[g] + Q.src, b, m->ph, (long long)lb});
        m->running.push_back(q);
    }
    m->waiting.swap(still);
    if (seeds.empty()) return B40C_OK;
    m->last_admit_ph = m->ph;
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
    // One launch for every admitted query: a row of workgroups per query
    // (at most 65535 rows per launch), each row filling its label slot.
    const MsArgs a = ms_args(m);
    const unsigned gx = (unsigned)std::min<int64_t>((max_vg + 255) / 256, 1024);
    for (int r0 = 0; r0 < cnt; r0 += 65535) {
        const int rows = std::min(cnt - r0, 65535);
        hipLaunchKernelGGL(k_ms_admit, dim3(gx, (unsigned)rows), dim3(256), 0, m->stream, a, m->ph,
                           m->seed_buf.as<MsSeed>() + r0, rows);
        B40C_HIP(hipGetLastError());
    }
    return B40C_OK;
}

// Retires the queries the device marked finished (k_ms_direction wrote their
// level counts) in the levels just run.
b40c_status ms_retire(b40c_msbfs* m) {
    std::vector<int32_t> keep;
    std::map<int32_t, unsigned long long> masks;  // graph -> retired slots
    for (int32_t q : m->running) {
        MsQuery& Q = m->queries[q];
        const unsigned long long bit = 1ull << Q.bit;
        const int32_t levels = __atomic_load_n(&m->h_done[q], __ATOMIC_ACQUIRE);
        if (levels <= 0) {
            keep.push_back(q);
            continue;
        }
        Q.levels = levels;
        Q.state = 2;
        m->completed.push_back(q);
        m->h_active[Q.graph] &= ~bit;
        masks[Q.graph] |= bit;
    }
    m->running.swap(keep);
    if (masks.empty()) return B40C_OK;
    // One launch for all graphs with retired slots, a row of workgroups per
    // graph.  Pinned staging: the step synchronized the stream, so the
    // previous level's release has completed.
    const int cnt = (int)masks.size();
    if (m->h_rel_cap < cnt) {
        if (m->h_rel) (void)hipHostFree(m->h_rel);
        m->h_rel = nullptr;
        m->h_rel_cap = 0;
        B40C_HIP(hipHostMalloc((void**)&m->h_rel, sizeof(MsRelease) * (size_t)std::max(cnt, 256), 0));
        m->h_rel_cap = std::max(cnt, 256);
    }
    int i = 0;
    int64_t max_range = 1;
    for (const auto& gm : masks) {
        m->h_rel[i++] = MsRelease{gm.first, 0, gm.second};
        max_range = std::max<int64_t>(max_range, m->mg->vbase[gm.first + 1] - m->mg->vbase[gm.first]);
    }
    B40C_TRY(m->rel_buf.reserve(sizeof(MsRelease) * (size_t)m->h_rel_cap));
    B40C_HIP(hipMemcpyAsync(m->rel_buf.ptr, m->h_rel, sizeof(MsRelease) * cnt, hipMemcpyHostToDevice, m->stream));
    const MsArgs a = ms_args(m);
    const unsigned gx = (unsigned)std::min<int64_t>((max_range + 255) / 256, 1024);
    for (int r0 = 0; r0 < cnt; r0 += 65535) {
        const int rows = std::min(cnt - r0, 65535);
        hipLaunchKernelGGL(k_ms_release, dim3(gx, (unsigned)rows), dim3(256), 0, m->stream, a,
                           m->rel_buf.as<MsRelease>() + r0, rows);
        B40C_HIP(hipGetLastError());
    }
    return B40C_OK;
}

// The six launches of one level (phase argument `ph`: absolute, or the
// level's index within a captured replay when a.dph is set).
void ms_launch_level(b40c_msbfs* m, const MsArgs& a, int ph) {
    hipLaunchKernelGGL(k_ms_push, dim3(m->grid), dim3(BFS_THREADS), 0, m->stream, a, ph);
    hipLaunchKernelGGL(k_ms_heavy, dim3(m->grid), dim3(BFS_THREADS), 0, m->stream, a, ph);
    hipLaunchKernelGGL(k_ms_pull_narrow, dim3(2 * m->grid), dim3(BFS_THREADS), 0, m->stream, a, ph);
    hipLaunchKernelGGL(k_ms_pull, dim3(2 * m->grid), dim3(BFS_THREADS), 0, m->stream, a, ph);
    hipLaunchKernelGGL(k_ms_count, dim3(m->grid), dim3(BFS_THREADS), 0, m->stream, a, ph);
    hipLaunchKernelGGL(k_ms_direction, dim3((m->G + BFS_WAVES - 1) / BFS_WAVES), dim3(BFS_THREADS), 0, m->stream, a,
                       ph);
}

// One k_ms_fused launch of `levels` levels: the largest co-resident grid
// (occupancy API) capped at the level kernels' push grid, a plain launch (no
// cooperative-launch path).
b40c_status ms_launch_fused(b40c_msbfs* m, int levels) {
    DeviceInfo info;
    B40C_TRY(device_info(info));
    int per_cu = 0;
    B40C_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_ms_fused, BFS_THREADS, 0));
    if (per_cu < 1) return fail(B40C_ERR_UNSUPPORTED, "batched persistent kernel cannot be resident");
    const int grid = std::min({per_cu * info.num_cus, m->grid, kMsMaxFusedGrid});
    MsArgs a = ms_args(m);
    int ph0 = m->ph;
    void* args[] = {&a, &ph0, &levels};
    B40C_HIP(hipLaunchKernel((const void*)k_ms_fused, dim3((unsigned)grid), dim3(BFS_THREADS), args, 0, m->stream));
    return B40C_OK;
}

// Captures levels_per_replay levels plus k_ms_advance into one hipGraph
// (kept while the launch shape and direction knobs are unchanged; every
// buffer is fixed for the engine's life).  No memset nodes: counters and
// masks are cleared by the level kernels themselves.
b40c_status ms_ensure_loop(b40c_msbfs* m) {
    const int k = m->levels_per_replay;
    if (m->loop_exec && m->loop_levels == k && m->loop_grid == m->grid && m->loop_alpha == m->alpha &&
        m->loop_beta == m->beta)
        return B40C_OK;
    m->release_loop();
    MsArgs a = ms_args(m);
    a.dph = m->dph.as<int32_t>();
    B40C_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < k; ++j) ms_launch_level(m, a, j);
    hipLaunchKernelGGL(k_ms_advance, dim3(1), dim3(64), 0, m->stream, a, k);
    const hipError_t el = hipGetLastError();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(m->stream, &graph);
    if (el != hipSuccess || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(B40C_ERR_HIP, "batched level loop capture failed: %s",
                    hipGetErrorString(el != hipSuccess ? el : ec));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(B40C_ERR_HIP, "batched level loop: hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    }
    m->loop_graph = graph;
    m->loop_exec = exec;
    m->loop_levels = k;
    m->loop_grid = m->grid;
    m->loop_alpha = m->alpha;
    m->loop_beta = m->beta;
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
    if (s == B40C_OK) s = mg->d_ebase.alloc(((size_t)ngraphs + 1) * 8);
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
    if (e == hipSuccess) e = hipMemcpy(mg->d_ebase.ptr, mg->ebase.data(), mg->ebase.size() * 8, hipMemcpyHostToDevice);
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
    if (mg->refs.fetch_sub(1) != 1) return;
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
    mg->refs.fetch_add(1);
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
    if (e == hipSuccess) e = hipHostMalloc((void**)&m->h_ctr, sizeof(MsCounters), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&m->h_ctr_dev, m->h_ctr, 0);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&m->h_done, (size_t)max_queries * sizeof(int32_t), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&m->h_done_dev, m->h_done, 0);
    if (e != hipSuccess) s = fail(B40C_ERR_HIP, "b40c_msbfs_create: %s", hipGetErrorString(e));
    const size_t N = (size_t)m->n;
    if (s == B40C_OK) s = m->visited.alloc(N * 8);
    for (int i = 0; i < 3 && s == B40C_OK; ++i) s = m->front[i].alloc(N * 8);
    if (s == B40C_OK) s = m->active.alloc((size_t)m->G * 8);
    if (s == B40C_OK) s = m->bu.alloc((size_t)m->G * 8);
    if (s == B40C_OK) s = m->dph.alloc(sizeof(uint32_t));
    if (s == B40C_OK) s = m->bar_arrive.alloc((size_t)kMsMaxFusedGrid * sizeof(uint32_t));
    if (s == B40C_OK) s = m->cnt_n.alloc((size_t)m->G * 64 * 4);
    if (s == B40C_OK) s = m->cnt_m.alloc((size_t)m->G * 64 * 8);
    if (s == B40C_OK) s = m->nf_prev.alloc((size_t)m->G * 64 * 4);
    if (s == B40C_OK) s = m->m_seen.alloc((size_t)m->G * 64 * 8);
    if (s == B40C_OK) s = m->bitq.alloc((size_t)m->G * 64 * 4);
    if (s == B40C_OK) s = m->lbase.alloc((size_t)max_queries * 8);
    if (s == B40C_OK) s = m->qstart.alloc((size_t)max_queries * 4);
    if (s == B40C_OK) s = m->depth.alloc((size_t)m->arena_elems * 4);
    if (s == B40C_OK) s = m->pred.alloc((size_t)m->arena_elems * 4);
    for (int i = 0; i < 3 && s == B40C_OK; ++i) s = m->queue[i].alloc(N * 4);
    if (s == B40C_OK) s = m->bu_list.alloc((size_t)m->G * 4 * sizeof(int2));
    if (s == B40C_OK) s = m->counters.alloc(sizeof(MsCounters));
    m->heavy_cap = (uint32_t)m->n;
    m->chunk_cap = (uint32_t)std::min<int64_t>(mg->packed->edges / HEAVY_CHUNK + (int64_t)m->n + 64, 0x7FFFFFFFll);
    if (s == B40C_OK) s = m->heavy.alloc((size_t)m->heavy_cap * sizeof(int2));
    if (s == B40C_OK) s = m->chunk_map.alloc((size_t)m->chunk_cap * 4);
    if (s == B40C_OK) s = m->fbm.alloc(((size_t)m->n + 31) / 32 * 4 * 2 + 4);
    if (s != B40C_OK) {
        b40c_msbfs_destroy(m);
        return s;
    }
    m->arena.init(m->arena_elems);
    m->h_active.assign(m->G, 0ull);
    std::memset(m->h_ctr, 0, sizeof(MsCounters));
    std::memset(m->h_done, 0, (size_t)max_queries * sizeof(int32_t));
    B40C_HIP(hipMemsetAsync(m->visited.ptr, 0, m->visited.bytes, m->stream));
    for (int i = 0; i < 3; ++i) B40C_HIP(hipMemsetAsync(m->front[i].ptr, 0, m->front[i].bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->dph.ptr, 0, m->dph.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->bar_arrive.ptr, 0, m->bar_arrive.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->bu.ptr, 0, m->bu.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->fbm.ptr, 0, m->fbm.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->cnt_n.ptr, 0, m->cnt_n.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->cnt_m.ptr, 0, m->cnt_m.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->nf_prev.ptr, 0, m->nf_prev.bytes, m->stream));
    B40C_HIP(hipMemsetAsync(m->m_seen.ptr, 0, m->m_seen.bytes, m->stream));
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

b40c_status b40c_msbfs_set_admission(b40c_msbfs* m, int32_t cohort_gap) {
    if (!m || cohort_gap < 1) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_set_admission: bad arguments");
    m->cohort_gap = cohort_gap;
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
    // Ids are recycled once a query's result has been taken, so max_queries
    // bounds the queries submitted and not yet collected, not the total.
    int32_t q = -1;
    if (!m->free_ids.empty()) {
        q = m->free_ids.back();
        m->free_ids.pop_back();
    } else if ((int32_t)m->queries.size() < m->qmax) {
        m->queries.emplace_back();
        q = (int32_t)m->queries.size() - 1;
    } else {
        return fail(B40C_ERR_ADMISSION, "b40c_msbfs_submit: %d queries submitted and not yet collected", m->qmax);
    }
    MsQuery Q;
    Q.graph = graph;
    Q.src = src;
    m->queries[q] = Q;
    m->waiting.push_back(q);
    *query_id = q;
    return B40C_OK;
}

// One level boundary: admit, run one level (or one captured replay of
// levels_per_replay levels) for every query in flight, retire.
b40c_status b40c_msbfs_step(b40c_msbfs* m, int32_t* running, int32_t* waiting) {
    if (!m) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_step: null");
    B40C_TRY(ms_admit(m));
    if (!m->running.empty()) {
        if (m->bar_dirty) {
            // A persistent launch's barrier wait timed out: clear the barrier
            // state (every form publishes its abort flag).
            B40C_HIP(hipMemsetAsync(m->bar_arrive.ptr, 0, m->bar_arrive.bytes, m->stream));
            B40C_HIP(hipMemsetAsync(&m->counters.as<MsCounters>()->bar_release, 0, sizeof(unsigned long long),
                                    m->stream));
            B40C_HIP(hipMemsetAsync(&m->counters.as<MsCounters>()->bar_abort, 0, sizeof(uint32_t), m->stream));
            m->bar_dirty = false;
        }
        int levels = 1;
        if (m->persistent) {
            levels = std::max(1, m->levels_per_replay);
            B40C_TRY(ms_launch_fused(m, levels));
        } else if (m->levels_per_replay > 0) {
            levels = m->levels_per_replay;
            MsArgs a = ms_args(m);
            a.dph = m->dph.as<int32_t>();
            if (m->dev_ph != m->ph) {
                hipLaunchKernelGGL(k_ms_set_phase, dim3(1), dim3(64), 0, m->stream, a, m->ph);
                B40C_HIP(hipGetLastError());
                m->dev_ph = m->ph;
            }
            B40C_TRY(ms_ensure_loop(m));
            B40C_HIP(hipGraphLaunch(m->loop_exec, m->stream));
            m->dev_ph += levels;
        } else {
            const MsArgs a = ms_args(m);
            ms_launch_level(m, a, m->ph);
            B40C_HIP(hipGetLastError());
        }
        // k_ms_direction has written h_ctr's flags and the finished
        // queries' done words.
        B40C_HIP(hipStreamSynchronize(m->stream));
        if (m->h_ctr->bar_abort) {
            m->bar_dirty = true;
            return fail(B40C_ERR_HIP, "batched traversal: a grid-barrier wait timed out (workgroups not co-resident)");
        }
        if (m->h_ctr->overflow || m->h_ctr->corrupt)
            return fail(B40C_ERR_HIP, "batched traversal: frontier overflow or corrupt frontier");
        B40C_TRY(ms_retire(m));
        m->ph += levels;
    }
    if (running) *running = (int32_t)m->running.size();
    if (waiting) *waiting = (int32_t)m->waiting.size();
    return B40C_OK;
}

b40c_status b40c_msbfs_set_levels_per_step(b40c_msbfs* m, int32_t levels) {
    if (!m || levels < 0 || levels > 64)
        return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_set_levels_per_step: bad arguments");
    m->levels_per_replay = levels;
    return B40C_OK;
}

b40c_status b40c_msbfs_set_persistent(b40c_msbfs* m, int32_t on) {
    if (!m || (on != 0 && on != 1)) return fail(B40C_ERR_INVALID_ARG, "b40c_msbfs_set_persistent: bad arguments");
    m->persistent = on != 0;
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
    // Without copies there is nothing to wait for: the slot's labels are no
    // longer written (the query retired), and whoever reuses the memory next
    // is ordered after everything already on the stream.
    if (depth || pred) B40C_HIP(hipStreamSynchronize(m->stream));
    if (search_depth) *search_depth = Q.levels;
    m->arena.release(Q.lbase, vg);
    Q.state = 3;
    m->free_ids.push_back(q);
    return B40C_OK;
}

void* b40c_msbfs_stream(b40c_msbfs* m) { return m ? (void*)m->stream : nullptr; }

void b40c_msbfs_destroy(b40c_msbfs* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->visited.release();
    for (int i = 0; i < 3; ++i) m->front[i].release();
    m->active.release();
    m->bu.release();
    m->release_loop();
    m->dph.release();
    m->bar_arrive.release();
    m->cnt_n.release();
    m->cnt_m.release();
    m->nf_prev.release();
    m->m_seen.release();
    m->bitq.release();
    m->lbase.release();
    m->qstart.release();
    m->depth.release();
    m->pred.release();
    for (int i = 0; i < 3; ++i) m->queue[i].release();
    m->bu_list.release();
    m->heavy.release();
    m->chunk_map.release();
    m->fbm.release();
    m->counters.release();
    m->seed_buf.release();
    m->rel_buf.release();
    if (m->h_seeds) (void)hipHostFree(m->h_seeds);
    if (m->h_rel) (void)hipHostFree(m->h_rel);
    if (m->h_ctr) (void)hipHostFree(m->h_ctr);
    if (m->h_done) (void)hipHostFree(m->h_done);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    b40c_multigraph_destroy(m->mg);
    delete m;
}

}  // extern "C"
