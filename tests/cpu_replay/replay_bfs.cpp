// CPU replay of the BFS kernels (test infrastructure only).
//
// Compiles the library's kernel bodies (back40computing_amd/csrc/
// bfs_kernels.hpp) for the host against the device-API emulation in
// include/hip/hip_runtime.h and drives them with the host loops of the
// enactors in bfs_host.inc (iterative level loops, the persistent kernels
// and hybrid's persistent/two-phase hand-off).  Built with AddressSanitizer:
// every device buffer is its own allocation of exactly the size the library
// allocates, buffers the library leaves uninitialised are filled with a
// garbage pattern, and every workgroup's LDS block starts with a garbage
// pattern too (LDS keeps whatever an earlier kernel left), so an index
// outside a buffer or a read of LDS that the kernel never wrote shows up as a
// sanitizer report or as wrong labels.
//
// usage: replay_bfs GRAPH SRC STRATEGY MODE GRID CTA_TH WAVE_TH GRID_TH SAT_QUIT QUEUE_SIZING LDS_FILL OUT
//        replay_bfs --gbarrier GRID ITERATIONS LIGHT
//   GRAPH: int32 n, int64 m, int32 offsets[n+1], int32 cols[m] (little endian)
//   STRATEGY/MODE: the library's numbering (0 EC, 1 CE, 2 TP, 3 hybrid, 4 persistent; 0 default, 1 fused, 2 iterative)
//   LDS_FILL: byte value every LDS block starts with
//   OUT: int32 status, int32 search_depth, int32 depth[n], int32 pred[n]
//   status: 0 ok, 1 queue overflow, 5 corrupt frontier id or barrier abort
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "bfs_kernels.hpp"

using namespace b40c;

namespace {

constexpr int kMaxFusedGrid = 4096;
constexpr int kStrEC = 0, kStrCE = 1, kStrTP = 2, kStrHybrid = 3, kStrPersistent = 4;
constexpr int kModeDefault = 0, kModeFused = 1, kModeIterative = 2;

// One device buffer: exactly `bytes` long, so AddressSanitizer reports any
// access past it.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void alloc(size_t n, int fill) {
        bytes = n;
        // 256-byte aligned like hipMalloc's blocks (DevCounters needs it).
        if (posix_memalign(&p, 256, n ? n : 1) != 0) {
            std::fprintf(stderr, "replay: allocation of %zu bytes failed\n", n);
            std::exit(2);
        }
        std::memset(p, fill, n ? n : 1);
    }
    ~DevBuf() { std::free(p); }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

int g_lds_fill = 0x3C;

template <class LDS, class F>
void run(unsigned grid, F body) {
    std::vector<LDS*> lds(grid);
    for (auto& l : lds) {
        l = static_cast<LDS*>(std::malloc(sizeof(LDS)));
        std::memset((void*)l, g_lds_fill, sizeof(LDS));
    }
    emu::launch(grid, BFS_THREADS, [&] { body(*lds[blockIdx.x]); });
    for (auto* l : lds) std::free(l);
}

struct Replay {
    int32_t n = 0;
    int64_t m = 0;
    DevBuf offs, cols, depth, pred, visited, fbm[3], heavy, chunk_map, qk[2], qv[2], counters, bar;
    BfsArgs a{};
    DevCounters* ctr = nullptr;
    unsigned grid = 1;
    bool heavy_needed = false;

    void setup(const std::vector<int32_t>& h_off, const std::vector<int32_t>& h_cols, int frontier_type,
               double qs) {
        n = (int32_t)h_off.size() - 1;
        m = (int64_t)h_cols.size();
        offs.alloc(h_off.size() * 4, 0);
        std::memcpy(offs.p, h_off.data(), h_off.size() * 4);
        cols.alloc(h_cols.size() * 4, 0);
        std::memcpy(cols.p, h_cols.data(), h_cols.size() * 4);
        const size_t words = ((size_t)n + 31) / 32;
        depth.alloc((size_t)n * 4, 0xA5);
        pred.alloc((size_t)n * 4, 0xA5);
        visited.alloc(words * 4, 0xA5);
        for (auto& f : fbm) f.alloc(words * 4, 0xA5);
        const uint32_t heavy_cap = (uint32_t)n;
        const uint32_t chunk_cap = (uint32_t)(m / HEAVY_CHUNK + n + 64);
        heavy.alloc((size_t)heavy_cap * sizeof(int2), 0xA5);
        chunk_map.alloc((size_t)chunk_cap * 4, 0xA5);
        // Queue sizes per frontier type, as b40c_problem_reset sizes them.
        const double dn = n, dm = m > 0 ? (double)m : 1.0;
        size_t want[2], wantv[2];
        if (frontier_type == 0) {
            want[0] = want[1] = (size_t)(dn * qs);
            wantv[0] = wantv[1] = 0;
        } else if (frontier_type == 1) {
            want[0] = want[1] = (size_t)(dm * qs);
            wantv[0] = wantv[1] = want[0];
        } else {
            want[0] = (size_t)(dn * qs);
            want[1] = (size_t)(dm * qs);
            wantv[0] = 0;
            wantv[1] = want[1];
        }
        for (int i = 0; i < 2; ++i) {
            if (want[i] < 1) want[i] = 1;
            qk[i].alloc(want[i] * 4, 0xA5);
            if (wantv[i]) qv[i].alloc(wantv[i] * 4, 0xA5);
            a.qk[i] = qk[i].as<int32_t>();
            a.qv[i] = wantv[i] ? qv[i].as<int32_t>() : nullptr;
            a.qcap[i] = (uint32_t)want[i];
        }
        counters.alloc(sizeof(DevCounters), 0);
        bar.alloc((size_t)kMaxFusedGrid * 4, 0);
        ctr = counters.as<DevCounters>();
        a.offs = offs.as<int32_t>();
        a.cols = cols.as<int32_t>();
        a.depth = depth.as<int32_t>();
        a.pred = pred.as<int32_t>();
        a.visited = visited.as<uint32_t>();
        for (int i = 0; i < 3; ++i) a.fbm[i] = fbm[i].as<uint32_t>();
        a.heavy = heavy.as<int2>();
        a.chunk_map = chunk_map.as<int32_t>();
        a.heavy_cap = heavy_cap;
        a.chunk_cap = chunk_cap;
        a.ctr = ctr;
        a.bar_arrive = bar.as<uint32_t>();
        a.hctr = nullptr;
        a.max_depth = &ctr->max_depth;
        a.wg_stats = nullptr;
        a.n = n;
        a.words = (int32_t)words;
        a.symmetric = 0;
        a.do_alpha = 14;
        a.do_beta = 24;
        a.iso = nullptr;
        a.dev_base = 0;
        a.edges = m;
    }

    void seed(int ph, int32_t src, int mode) {
        const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(((int64_t)n + 255) / 256, 8));
        emu::launch(g, 256, [&] { reset_seed_body(a, ph, src, mode); });
    }
    void phase(int kind, int ph, int32_t depth_arg) {
        run<LevelShared>(grid, [&](LevelShared& s) {
            switch (kind) {
                case PK_TD: phase_body<PK_TD>(a, ph, depth_arg, 0, s); break;
                case PK_CE: phase_body<PK_CE>(a, ph, depth_arg, 0, s); break;
                case PK_TP_CONTRACT: phase_body<PK_TP_CONTRACT>(a, ph, depth_arg, 0, s); break;
                case PK_TP_EXPAND: phase_body<PK_TP_EXPAND>(a, ph, depth_arg, 0, s); break;
                case PK_TP_FILTER: phase_body<PK_TP_FILTER>(a, ph, depth_arg, 0, s); break;
            }
        });
    }
    void heavy_step(int kind, int ph, int32_t depth_next) {
        if (!heavy_needed) return;
        if (kind == HK_TD)
            run<HeavySharedOf<HK_TD>::type>(grid, [&](HeavySharedOf<HK_TD>::type& s) {
                heavy_body<HK_TD>(a, ph, depth_next, 0, s);
            });
        else
            run<HeavySharedOf<HK_PAIRS>::type>(grid, [&](HeavySharedOf<HK_PAIRS>::type& s) {
                heavy_body<HK_PAIRS>(a, ph, depth_next, 0, s);
            });
    }
    template <int KIND>
    struct FusedLds {
        LevelShared s;
        typename FusedHeavy<KIND>::shared hs;
    };
    void fused(int kind, int ph, int32_t depth_arg, uint32_t sat) {
        BfsArgs fa = a;
        if (!heavy_needed) fa.grid_th = INT_MAX;
        // As the library: agent-scope frontier accesses and the light
        // barrier in every phase.
        fa.persistent = 1;
        fa.coh_in = 1;
        fa.coh_out = 1;
        if (kind == FK_TD)
            run<FusedLds<FK_TD>>(grid, [&](FusedLds<FK_TD>& l) {
                fused_body<FK_TD>(fa, ph, depth_arg, 1 << 30, sat, l.s, l.hs);
            });
        if (kind == FK_CE)
            run<FusedLds<FK_CE>>(grid, [&](FusedLds<FK_CE>& l) {
                fused_body<FK_CE>(fa, ph, depth_arg, 1 << 30, sat, l.s, l.hs);
            });
        if (kind == FK_TP)
            run<FusedLds<FK_TP>>(grid, [&](FusedLds<FK_TP>& l) {
                fused_body<FK_TP>(fa, ph, depth_arg, 1 << 30, sat, l.s, l.hs);
            });
    }
    bool stop(int ph) const { return ctr->overflow || ctr->corrupt || ctr->bar_abort || ctr->qlen[ph & 3] == 0; }
};

std::vector<char> read_file(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        std::perror(path);
        std::exit(2);
    }
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    std::fclose(f);
    return buf;
}

}  // namespace

// Grid-barrier visibility check (the library's gbarrier_check_body) on
// `grid` workgroups for `iters` iterations; prints the mismatch count.
int gbarrier_check(unsigned grid, int iters, int light) {
    const int words = 256;
    DevBuf data, bad, arrive, rel, abort_flag;
    data.alloc((size_t)grid * words * 4, 0);
    bad.alloc(8, 0);
    arrive.alloc((size_t)grid * 4, 0);
    rel.alloc(8, 0);
    abort_flag.alloc(4, 0);
    const GridBarrier bar{arrive.as<uint32_t>(), rel.as<unsigned long long>(), abort_flag.as<uint32_t>(),
                          (uint32_t)light};
    for (int round = 0; round < 2; ++round)  // a second launch continues the epochs of the first
        run<uint32_t>(grid, [&](uint32_t& s_word) {
            gbarrier_check_body(bar, data.as<uint32_t>(), words, iters, bad.as<unsigned long long>(), &s_word);
        });
    std::printf("mismatches %llu aborted %u epoch %llu\n", *bad.as<unsigned long long>(), *abort_flag.as<uint32_t>(),
                *rel.as<unsigned long long>() & 0xFFFFFFFFull);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 5 && std::string(argv[1]) == "--gbarrier")
        return gbarrier_check((unsigned)std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    if (argc != 13) {
        std::fprintf(stderr,
                     "usage: %s GRAPH SRC STRATEGY MODE GRID CTA_TH WAVE_TH GRID_TH SAT_QUIT QUEUE_SIZING LDS_FILL "
                     "OUT\n",
                     argv[0]);
        return 2;
    }
    const std::vector<char> gbuf = read_file(argv[1]);
    int32_t n = 0;
    int64_t m = 0;
    std::memcpy(&n, gbuf.data(), 4);
    std::memcpy(&m, gbuf.data() + 4, 8);
    if (gbuf.size() != 12 + 4 * ((size_t)n + 1) + 4 * (size_t)m) {
        std::fprintf(stderr, "replay: graph file has the wrong size\n");
        return 2;
    }
    std::vector<int32_t> off((size_t)n + 1), cols((size_t)m);
    std::memcpy(off.data(), gbuf.data() + 12, 4 * ((size_t)n + 1));
    std::memcpy(cols.data(), gbuf.data() + 12 + 4 * ((size_t)n + 1), 4 * (size_t)m);
    const int32_t src = std::atoi(argv[2]);
    const int strategy = std::atoi(argv[3]);
    const int mode = std::atoi(argv[4]);
    const unsigned grid = (unsigned)std::atoi(argv[5]);
    const int cta_th = std::atoi(argv[6]), wave_th = std::atoi(argv[7]), grid_th = std::atoi(argv[8]);
    const int sat_quit = std::atoi(argv[9]);
    const double qs = std::atof(argv[10]);
    g_lds_fill = std::atoi(argv[11]) & 0xFF;
    if (src < 0 || src >= n || grid < 1 || grid > (unsigned)kMaxFusedGrid) {
        std::fprintf(stderr, "replay: bad source or grid\n");
        return 2;
    }
    const int frontier_type = (strategy == kStrEC || strategy == kStrPersistent) ? 0 : strategy == kStrTP ? 2 : 1;
    Replay r;
    r.setup(off, cols, frontier_type, qs);
    r.grid = grid;
    r.a.cta_th = std::max(1, cta_th);
    r.a.wave_th = std::max(1, wave_th);
    r.a.grid_th = std::max(1, grid_th);
    r.a.td_grid = (int32_t)grid;
    int32_t max_deg = 0;
    for (int32_t v = 0; v < n; ++v) max_deg = std::max(max_deg, off[v + 1] - off[v]);
    r.heavy_needed = max_deg >= r.a.grid_th;
    const bool fused = (mode == kModeFused) || strategy == kStrPersistent ||
                       (mode == kModeDefault && (strategy == kStrEC || strategy == kStrCE));
    int ph = 0;
    int32_t depth = 0;
    if (strategy == kStrEC || strategy == kStrPersistent) {
        r.seed(0, src, SEED_VERTEX);
        if (fused && !(strategy == kStrPersistent && mode == kModeIterative)) {
            r.fused(FK_TD, 0, 0, 0);
        } else {
            do {
                r.phase(PK_TD, ph, ph + 1);
                r.heavy_step(HK_TD, ph, ph + 1);
                ++ph;
            } while (!r.stop(ph));
        }
    } else if (strategy == kStrCE || (strategy == kStrHybrid && mode != kModeDefault)) {
        r.seed(0, src, SEED_EDGE);
        if (mode == kModeFused || (strategy == kStrCE && mode == kModeDefault)) {
            r.fused(FK_CE, 0, 0, 0);
        } else {
            do {
                r.phase(PK_CE, ph, ph);
                r.heavy_step(HK_PAIRS, ph, ph);
                ++ph;
            } while (!r.stop(ph));
        }
    } else if (strategy == kStrTP) {
        ph = 1;
        r.seed(1, src, SEED_EDGE);
        if (mode == kModeFused) {
            r.fused(FK_TP, 1, 0, 0);
        } else {
            do {
                r.phase(PK_TP_CONTRACT, ph, depth);
                r.phase(PK_TP_EXPAND, ph + 1, depth);
                r.heavy_step(HK_PAIRS, ph + 1, depth);
                ph += 2;
                ++depth;
            } while (!r.stop(ph));
        }
    } else if (strategy == kStrHybrid) {
        // enact_hybrid: persistent contract-expand while the frontier is at
        // most grid x saturation_quit, filter -> contract -> expand steps on
        // wider levels, back when the filtered frontier is small again.
        r.seed(0, src, SEED_EDGE);
        const uint32_t sat = (uint32_t)std::max<int64_t>(1, (int64_t)grid * std::max(1, sat_quit));
        while (true) {
            r.fused(FK_CE, ph, depth, sat);
            const uint32_t levels = r.ctr->levels_run;
            ph = (int)r.ctr->last_level;
            depth += (int32_t)levels;
            if (r.ctr->overflow || r.ctr->corrupt || r.ctr->bar_abort || !r.ctr->saturated) break;
            while (true) {
                r.phase(PK_TP_FILTER, ph, depth);
                r.phase(PK_TP_CONTRACT, ph + 1, depth);
                r.phase(PK_TP_EXPAND, ph + 2, depth);
                r.heavy_step(HK_PAIRS, ph + 2, depth);
                ph += 3;
                ++depth;
                if (r.ctr->overflow || r.ctr->corrupt) break;
                const uint32_t q = r.ctr->qlen[ph & 3];
                if (q == 0 || q <= sat) break;
            }
            if (r.stop(ph)) break;
        }
    } else {
        std::fprintf(stderr, "replay: strategy %d is not replayed\n", strategy);
        return 2;
    }
    int32_t status = 0;
    if (r.ctr->corrupt || r.ctr->bar_abort) status = 5;
    else if (r.ctr->overflow) status = 1;
    const int32_t search_depth = (int32_t)r.ctr->max_depth + 1;
    FILE* f = std::fopen(argv[12], "wb");
    if (!f) {
        std::perror(argv[12]);
        return 2;
    }
    std::fwrite(&status, 4, 1, f);
    std::fwrite(&search_depth, 4, 1, f);
    std::fwrite(r.depth.p, 4, (size_t)n, f);
    std::fwrite(r.pred.p, 4, (size_t)n, f);
    std::fclose(f);
    return 0;
}
