/*
 * C++ facade for the per-stage BFS microbenchmarks, with the reference's
 * class names:
 *   bfs::microbench::EnactorGatherLookup   test/bfs/microbench/enactor_gather_lookup.cuh:57-558
 *   bfs::microbench::EnactorCull           test/bfs/microbench/enactor_cull.cuh:57-493
 * EnactSearch runs the search from `src` over a device graph (the graph a
 * CsrProblem was built from), keeps the per-iteration records and optionally
 * copies the depth labels to a host array; GetStatistics reports the
 * reference's figures.  EnactorGatherLookup's total_queued is the sum of the
 * expansion queues (the ids entering the status lookups), EnactorCull's
 * total_unculled the sum of the ids left by the bench culls.  The duty
 * figures are 0: these enactors time every launch with events instead
 * (Levels()).
 */
#pragma once

#include <cstdint>
#include <vector>

#include "../b40c_mi355x.h"

namespace b40c {
namespace graph {
namespace bfs {
namespace microbench {

template <int MODE>
class MicrobenchEnactor {
public:
    b40c_status EnactSearch(const b40c_graph* graph, int32_t src, int32_t* h_depth = nullptr) {
        levels_.resize(levels_.empty() ? 256 : levels_.size());
        int32_t n = 0;
        b40c_status s = b40c_microbench_search(graph, src, MODE, h_depth, levels_.data(), (int32_t)levels_.size(), &n);
        if (s == B40C_OK && n > (int32_t)levels_.size()) {
            levels_.resize((size_t)n);
            s = b40c_microbench_search(graph, src, MODE, h_depth, levels_.data(), (int32_t)levels_.size(), &n);
        }
        levels_.resize(s == B40C_OK ? (size_t)n : 0);
        return s;
    }
    // Per-iteration records of the last search.
    const std::vector<b40c_mb_level>& Levels() const { return levels_; }
    // search_depth = deepest label + 1 (the enactors' convention).
    void GetStatistics(long long& total, int32_t& search_depth, double& expand_duty, double& compact_duty) const {
        total = 0;
        search_depth = 0;
        for (const b40c_mb_level& r : levels_) {
            if (r.iteration > 0) total += (MODE == B40C_MB_CULL) ? r.local_pass : r.lookup_in;
            if (r.discovered > 0) search_depth = r.iteration + 1;
        }
        expand_duty = 0.0;
        compact_duty = 0.0;
    }

private:
    std::vector<b40c_mb_level> levels_;
};

typedef MicrobenchEnactor<B40C_MB_GATHER_LOOKUP> EnactorGatherLookup;
typedef MicrobenchEnactor<B40C_MB_CULL> EnactorCull;

}  // namespace microbench
}  // namespace bfs
}  // namespace graph
}  // namespace b40c
