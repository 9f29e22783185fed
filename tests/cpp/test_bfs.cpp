// C++ parity driver in the shape of the reference's test/bfs/test_bfs.cu
// RunTests loop (test_bfs.cu:584-783): build a graph, then per source
// Reset -> timed EnactSearch -> GetStatistics -> ExtractResults -> validate
// against the serial CPU BFS (the oracle, test infrastructure) and the
// predecessor invariants (test_bfs.cu:387-438).  It drives the library only
// through the C++ facade (include/b40c/bfs.hpp), i.e. the way a
// reference-side caller would after relinking.
//
// usage: test_bfs rmat <scale> [edge_factor] | grid3d <width>
//                 | market <file> | metis <file> | dimacs <file>
//                 [--undirected] [--splice=<n>]
//                 [--strategy=0..5] [--mode=default|fused|iterative]
//                 [--i=<sources>] [--src=<vertex>] [--seed=<n>]
//                 [--queue-sizing=<f>] [--instrumented] [--v2]
//                 [--microbench=gather_lookup|cull]
// Prints one line per source and a summary; exit status 0 iff every source
// is CORRECT.  --microbench runs the per-stage microbenchmark search instead
// of an enactor, as the reference's microbench_bfs.cu / microbench_cull.cu
// drivers do (test/bfs/microbench/microbench_*.cu: the same driver over
// EnactorGatherLookup / EnactorCull): per source the queue lengths of every
// iteration, the stage times, the labels checked against the CPU BFS, and in
// cull mode the culling efficiency (microbench_cull.cu:313-331), through
// the facade's microbench::EnactorGatherLookup / EnactorCull
// (include/b40c/microbench.hpp).  As in the reference: --instrumented runs
// the instrumented enactors (average workgroup duty), each source's line
// carries total queued
// and redundant work ((queued - edges visited) / edges visited,
// test_bfs.cu:369-373), the summary the mean, deviation and cv of search
// depth, redundant work, duty and rate (test_bfs.cu:451-478), and --v2 prints
// the work histogram by depth of every source (test_bfs.cu:246-337, 560).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "b40c/bfs.hpp"
#include "b40c/graph_builder.hpp"
#include "b40c/microbench.hpp"

extern "C" {
int32_t orc_bfs(int32_t n, const int32_t* offsets, const int32_t* cols, int32_t src, int32_t* depth, int32_t* pred);
int64_t orc_check_pred(int32_t n, const int32_t* offsets, const int32_t* cols, const int32_t* ref_depth,
                       const int32_t* pred);
int64_t orc_edges_visited(int32_t n, const int32_t* offsets, const int32_t* labels, int64_t* nodes_visited);
int32_t orc_pick_sources(int32_t n, const int32_t* offsets, uint64_t seed, int32_t count, int32_t* out);
int32_t orc_histogram(int32_t n, const int32_t* offsets, const int32_t* cols, const int32_t* labels,
                      int32_t search_depth, int64_t* expanded, int64_t* unique_expanded, int64_t* discovered);
}

using namespace b40c::graph::bfs;
typedef CsrProblem<int, int, true> Problem;

namespace {

struct Options {
    std::string kind;
    int scale = 12, edge_factor = 16, width = 32;
    int strategy = B40C_HYBRID;
    std::string mode = "default";
    int iterations = 4;
    int src = -1;
    unsigned long long seed = 1;
    double queue_sizing = 1.3;
    bool instrumented = false;
    bool v2 = false;
    std::string file;
    bool undirected = false;  // file graphs: add a back edge per edge (test_bfs.cu:182-185)
    int splice = 0;
    std::string microbench;   // "", "gather_lookup" or "cull"
};

bool file_kind(const std::string& k) { return k == "market" || k == "metis" || k == "dimacs"; }

bool parse(int argc, char** argv, Options& o) {
    int pos = 0;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (!strncmp(a, "--strategy=", 11)) o.strategy = atoi(a + 11);
        else if (!strncmp(a, "--mode=", 7)) o.mode = a + 7;
        else if (!strncmp(a, "--i=", 4)) o.iterations = atoi(a + 4);
        else if (!strncmp(a, "--src=", 6)) o.src = atoi(a + 6);
        else if (!strncmp(a, "--seed=", 7)) o.seed = strtoull(a + 7, nullptr, 10);
        else if (!strncmp(a, "--queue-sizing=", 15)) o.queue_sizing = atof(a + 15);
        else if (!strcmp(a, "--undirected")) o.undirected = true;
        else if (!strcmp(a, "--instrumented")) o.instrumented = true;
        else if (!strcmp(a, "--v2")) o.v2 = true;
        else if (!strncmp(a, "--splice=", 9)) o.splice = atoi(a + 9);
        else if (!strncmp(a, "--microbench=", 13)) {
            o.microbench = a + 13;
            if (o.microbench != "gather_lookup" && o.microbench != "cull") return false;
        }
        else if (a[0] == '-') return false;
        else if (pos == 0) { o.kind = a; ++pos; }
        else if (o.kind == "rmat" && pos == 1) { o.scale = atoi(a); ++pos; }
        else if (o.kind == "rmat" && pos == 2) { o.edge_factor = atoi(a); ++pos; }
        else if (o.kind == "grid3d" && pos == 1) { o.width = atoi(a); ++pos; }
        else if (file_kind(o.kind) && pos == 1) { o.file = a; ++pos; }
        else return false;
    }
    return o.kind == "rmat" || o.kind == "grid3d" || (file_kind(o.kind) && !o.file.empty());
}

// Running mean and sample variance (Welford).
struct RunningStat {
    long long count = 0;
    double mean = 0.0, m2 = 0.0;
    void add(double x) {
        ++count;
        const double d = x - mean;
        mean += d / count;
        m2 += d * (x - mean);
    }
    double stddev() const { return count > 1 ? std::sqrt(m2 / (count - 1)) : 0.0; }
    void print(const char* name) const {
        printf("    [%s] u: %.3f, s: %.3f, cv: %.4f\n", name, mean, stddev(), mean != 0.0 ? stddev() / mean : 0.0);
    }
};

void print_histogram(const std::vector<int32_t>& offs, const std::vector<int32_t>& cols,
                     const std::vector<int32_t>& ref, int32_t search_depth) {
    const int32_t n = (int32_t)offs.size() - 1;
    std::vector<int64_t> ex(search_depth + 1), ue(search_depth + 1), dc(search_depth + 1);
    if (orc_histogram(n, offs.data(), cols.data(), ref.data(), search_depth, ex.data(), ue.data(), dc.data()) != 0)
        return;
    printf("Work Histogram:\nDepth, Expanded, Unique-Expanded, Discovered\n");
    for (int32_t d = 0; d < search_depth; ++d)
        printf("%d, %lld, %lld, %lld\n", d, (long long)ex[d], (long long)ue[d], (long long)dc[d]);
}

// The enactor classes are distinct types (as in the reference); run the
// source loop once per type.
template <typename Enactor>
int run(const Options& o, b40c_graph* g, const std::vector<int32_t>& offs, const std::vector<int32_t>& cols) {
    const int32_t n = (int32_t)offs.size() - 1;
    Problem problem;
    b40c_status s = problem.FromGraph(g);
    if (s != B40C_OK) {
        fprintf(stderr, "FromGraph: %s\n", b40c_last_error_string());
        return 2;
    }
    Enactor enactor(false);
    if (enactor.status() != B40C_OK) {
        fprintf(stderr, "enactor: %s\n", b40c_last_error_string());
        return 2;
    }
    std::vector<int32_t> sources;
    if (o.src >= 0) {
        sources.push_back(o.src);
    } else {
        sources.resize(o.iterations);
        sources.resize(orc_pick_sources(n, offs.data(), o.seed, o.iterations, sources.data()));
    }
    std::vector<int32_t> depth(n), pred(n), ref(n);
    int failures = 0;
    double total_ms = 0.0;
    long long total_edges = 0;
    RunningStat st_depth, st_redundant, st_duty, st_rate;
    for (int32_t src : sources) {
        s = problem.Reset(enactor.GetFrontierType(), o.queue_sizing);
        if (s != B40C_OK) {
            fprintf(stderr, "Reset: %s\n", b40c_last_error_string());
            return 2;
        }
        const auto t0 = std::chrono::steady_clock::now();
        if (o.mode == "fused") s = enactor.EnactFusedSearch(problem, src);
        else if (o.mode == "iterative") s = enactor.EnactIterativeSearch(problem, src);
        else s = enactor.EnactSearch(problem, src);
        long long queued = 0;
        int search_depth = 0;
        double duty = 0.0;
        if (s == B40C_OK) s = enactor.GetStatistics(queued, search_depth, duty);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (s != B40C_OK) {
            printf("src %d: status %d (%s)\n", src, s, b40c_last_error_string());
            ++failures;
            continue;
        }
        if (problem.ExtractResults(depth.data(), pred.data()) != B40C_OK) {
            printf("src %d: extract failed (%s)\n", src, b40c_last_error_string());
            ++failures;
            continue;
        }
        const int32_t ref_sd = orc_bfs(n, offs.data(), cols.data(), src, ref.data(), nullptr);
        const bool depth_ok = depth == ref;
        const int64_t bad_pred = orc_check_pred(n, offs.data(), cols.data(), ref.data(), pred.data());
        const bool ok = depth_ok && bad_pred < 0 && search_depth == ref_sd;
        int64_t nodes_visited = 0;
        const int64_t edges_visited = orc_edges_visited(n, offs.data(), ref.data(), &nodes_visited);
        total_ms += ms;
        total_edges += edges_visited;
        const double rate = ms > 0 ? edges_visited / (ms * 1000.0) : 0.0;
        const double redundant =
            (queued > 0 && edges_visited > 0) ? 100.0 * ((double)queued - edges_visited) / edges_visited : 0.0;
        printf("src %d: %s depth=%s pred=%s search_depth=%d (ref %d) nodes_visited=%lld edges_visited=%lld "
               "total_queued=%lld redundant_work=%.2f%% elapsed=%.3f ms rate=%.2f MiEdges/s",
               src, ok ? "CORRECT" : "INCORRECT", depth_ok ? "ok" : "MISMATCH", bad_pred < 0 ? "ok" : "BAD",
               search_depth, ref_sd, (long long)nodes_visited, (long long)edges_visited, queued, redundant, ms,
               rate);
        if (duty > 0) printf(" avg_duty=%.2f%%", duty * 100.0);
        printf("\n");
        if (o.v2) print_histogram(offs, cols, ref, ref_sd);
        st_depth.add(search_depth);
        st_redundant.add(redundant);
        st_duty.add(duty * 100.0);
        st_rate.add(rate);
        failures += ok ? 0 : 1;
    }
    if (st_rate.count > 0) {
        printf("  Summary after %lld test iterations (bias-corrected):\n", st_rate.count);
        st_depth.print("Search depth");
        st_redundant.print("redundant work %");
        if (st_duty.mean > 0) st_duty.print("Duty %");
        st_rate.print("Rate MiEdges/s");
    }
    printf("summary: strategy=%d mode=%s instrumented=%d sources=%zu failures=%d avg_elapsed=%.3f ms "
           "avg_rate=%.2f MiEdges/s\n",
           o.strategy, o.mode.c_str(), o.instrumented ? 1 : 0, sources.size(), failures,
           sources.empty() ? 0.0 : total_ms / sources.size(), total_ms > 0 ? total_edges / (total_ms * 1000.0) : 0.0);
    return failures ? 1 : 0;
}

// Per-stage microbenchmark search per source (the microbench drivers' loop):
// the reference's queue columns per iteration, the stage times, labels
// validated against the CPU BFS.  Elapsed = the real lookup, gather set-up
// and gather kernels (the bench passes are reported beside it).
template <typename MbEnactor>
int run_microbench(const Options& o, b40c_graph* g, const std::vector<int32_t>& offs,
                   const std::vector<int32_t>& cols) {
    const int32_t n = (int32_t)offs.size() - 1;
    const bool cull = o.microbench == "cull";
    MbEnactor enactor;
    std::vector<int32_t> sources;
    if (o.src >= 0) {
        sources.push_back(o.src);
    } else {
        sources.resize(o.iterations);
        sources.resize(orc_pick_sources(n, offs.data(), o.seed, o.iterations, sources.data()));
    }
    std::vector<int32_t> depth(n), ref(n);
    int failures = 0;
    RunningStat st_depth, st_cull, st_rate;
    for (int32_t src : sources) {
        const b40c_status s = enactor.EnactSearch(g, src, depth.data());
        const std::vector<b40c_mb_level>& lv = enactor.Levels();
        const int32_t nl = (int32_t)lv.size();
        if (s != B40C_OK) {
            printf("src %d: status %d (%s)\n", src, s, b40c_last_error_string());
            ++failures;
            continue;
        }
        printf(cull ? "Iteration, Bitmask cull, Label cull, Local cull, Compaction queue, Expansion queue, "
                      "bench lookup ms, lookup ms, offsets ms, bench gather ms, gather ms\n"
                    : "Iteration, Bitmask cull, Compaction queue, Expansion queue, "
                      "bench lookup ms, lookup ms, offsets ms, bench gather ms, gather ms\n");
        double ms = 0.0, bench_ms = 0.0;
        long long unculled = 0, total = 0;
        int32_t sd = 0;
        double expand_duty = 0.0, compact_duty = 0.0;
        enactor.GetStatistics(total, sd, expand_duty, compact_duty);
        for (int32_t i = 0; i < nl; ++i) {
            const b40c_mb_level& r = lv[i];
            if (cull)
                printf("%d, %lld, %lld, %lld, %lld, %lld, %.4f, %.4f, %.4f, %.4f, %.4f\n", r.iteration,
                       (long long)r.bitmask_pass, (long long)r.label_pass, (long long)r.local_pass,
                       (long long)r.discovered, (long long)r.gathered, r.bench_lookup_ms, r.lookup_ms, r.offsets_ms,
                       r.bench_gather_ms, r.gather_ms);
            else
                printf("%d, %lld, %lld, %lld, %.4f, %.4f, %.4f, %.4f, %.4f\n", r.iteration,
                       (long long)r.bitmask_pass, (long long)r.discovered, (long long)r.gathered, r.bench_lookup_ms,
                       r.lookup_ms, r.offsets_ms, r.bench_gather_ms, r.gather_ms);
            ms += r.lookup_ms + r.offsets_ms + r.gather_ms;
            bench_ms += r.bench_lookup_ms + r.bench_gather_ms;
            if (r.iteration > 0) unculled += cull ? r.local_pass : r.label_pass;
        }
        const int32_t ref_sd = orc_bfs(n, offs.data(), cols.data(), src, ref.data(), nullptr);
        const bool ok = depth == ref && sd == ref_sd;
        int64_t nodes_visited = 0;
        const int64_t edges_visited = orc_edges_visited(n, offs.data(), ref.data(), &nodes_visited);
        const double rate = ms > 0 ? edges_visited / (ms * 1000.0) : 0.0;
        // Share of the redundant lookups that the bench culls removed: every
        // vertex but the source needs one surviving copy, the other
        // edges_visited - (nodes_visited - 1) lookups are redundant
        // (microbench_cull.cu:327-331 divides by edges - nodes).
        const int64_t redundant = edges_visited - (nodes_visited - 1);
        const double eff = redundant > 0 ? 100.0 * double(edges_visited - unculled) / double(redundant) : 0.0;
        printf("src %d: %s depth=%s search_depth=%d (ref %d) nodes_visited=%lld edges_visited=%lld "
               "%s=%lld culling_efficiency=%.4f%% elapsed=%.3f ms bench=%.3f ms rate=%.2f MiEdges/s\n",
               src, ok ? "CORRECT" : "INCORRECT", depth == ref ? "ok" : "MISMATCH", sd, ref_sd,
               (long long)nodes_visited, (long long)edges_visited, cull ? "total_unculled" : "total_queued", total,
               eff, ms, bench_ms, rate);
        st_depth.add(sd);
        st_cull.add(eff);
        st_rate.add(rate);
        failures += ok ? 0 : 1;
    }
    if (st_rate.count > 0) {
        printf("  Summary after %lld test iterations (bias-corrected):\n", st_rate.count);
        st_depth.print("Search depth");
        st_cull.print("Culling efficiency %");
        st_rate.print("Rate MiEdges/s");
    }
    printf("summary: microbench=%s sources=%zu failures=%d\n", o.microbench.c_str(), sources.size(), failures);
    return failures ? 1 : 0;
}

template <bool INSTRUMENT>
int run_strategy(const Options& o, b40c_graph* g, const std::vector<int32_t>& offs, const std::vector<int32_t>& cols) {
    switch (o.strategy) {
        case B40C_EXPAND_CONTRACT: return run<EnactorExpandContract<INSTRUMENT>>(o, g, offs, cols);
        case B40C_CONTRACT_EXPAND: return run<EnactorContractExpand<INSTRUMENT>>(o, g, offs, cols);
        case B40C_TWO_PHASE: return run<EnactorTwoPhase<INSTRUMENT>>(o, g, offs, cols);
        case B40C_HYBRID: return run<EnactorHybrid<INSTRUMENT>>(o, g, offs, cols);
        case B40C_PERSISTENT: return run<EnactorPersistent<INSTRUMENT>>(o, g, offs, cols);
        case B40C_DIRECTION_OPTIMIZING: return run<EnactorDirectionOptimizing<INSTRUMENT>>(o, g, offs, cols);
    }
    fprintf(stderr, "unknown strategy %d\n", o.strategy);
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) {
        fprintf(stderr, "usage: %s rmat <scale> [edge_factor] | grid3d <width> | market <file> | metis <file> "
                        "| dimacs <file> [--undirected] [--splice=N] [--strategy=0..5] "
                        "[--mode=default|fused|iterative] [--i=N] [--src=V] [--seed=S] [--queue-sizing=F] "
                        "[--instrumented] [--v2] [--microbench=gather_lookup|cull]\n",
                argv[0]);
        return 2;
    }
    b40c_graph* g = nullptr;
    namespace builder = b40c::graph::builder;
    b40c_status s;
    if (o.kind == "rmat") s = b40c_graph_build_rmat(o.scale, o.edge_factor, 0.57, 0.19, 0.19, o.seed, &g);
    else if (o.kind == "grid3d") s = b40c_graph_build_lattice3d(o.width, o.seed, nullptr, &g);
    else if (o.kind == "market") s = builder::BuildMarketGraph(o.file.c_str(), &g, o.undirected);
    else if (o.kind == "metis") s = builder::BuildMetisGraph(o.file.c_str(), &g);
    else s = builder::BuildDimacsGraph(o.file.c_str(), &g, o.undirected, o.splice);
    if (s != B40C_OK) {
        fprintf(stderr, "graph build: %s\n", b40c_last_error_string());
        return 2;
    }
    int32_t n = 0;
    int64_t m = 0;
    b40c_graph_info(g, &n, &m);
    std::vector<int32_t> offs((size_t)n + 1), cols((size_t)(m > 0 ? m : 1));
    b40c_graph_copy_to_host(g, offs.data(), cols.data());
    printf("graph %s: nodes=%d edges=%lld\n", o.kind.c_str(), n, (long long)m);
    int rc = 2;
    if (o.microbench == "cull") rc = run_microbench<b40c::graph::bfs::microbench::EnactorCull>(o, g, offs, cols);
    else if (!o.microbench.empty())
        rc = run_microbench<b40c::graph::bfs::microbench::EnactorGatherLookup>(o, g, offs, cols);
    else rc = o.instrumented ? run_strategy<true>(o, g, offs, cols) : run_strategy<false>(o, g, offs, cols);
    b40c_graph_destroy(g);
    return rc;
}
