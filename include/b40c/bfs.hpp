/*
 * C++ facade over the b40c_mi355x C ABI with the reference's class and
 * method names, so test_bfs-style callers port by relinking against
 * libb40c_mi355x.so.  Header-only, no HIP or CUDA headers: every call goes
 * through include/b40c_mi355x.h.
 *
 * Reference interfaces mirrored (b40c tree, file:line):
 *   CsrProblem<VertexId, SizeT, MARK_PREDECESSORS>    b40c/graph/bfs/csr_problem.cuh:68-173
 *     FromHostProblem                                 csr_problem.cuh:328-403
 *     Reset(FrontierType, queue_sizing)               csr_problem.cuh:547-750
 *     ExtractResults(h_labels)                        csr_problem.cuh:255-322
 *   EnactorContractExpand<INSTRUMENT>                 b40c/graph/bfs/enactor_contract_expand.cuh:60-598
 *     EnactFusedSearch / EnactIterativeSearch         enactor_contract_expand.cuh:338-342, 525-529
 *     GetStatistics                                   enactor_contract_expand.cuh:236-250
 *   EnactorExpandContract<INSTRUMENT>                 b40c/graph/bfs/enactor_expand_contract.cuh:58-526
 *   EnactorTwoPhase<INSTRUMENT>                       b40c/graph/bfs/enactor_two_phase.cuh:67-766
 *   EnactorHybrid<INSTRUMENT>::EnactSearch            b40c/graph/bfs/enactor_hybrid.cuh:559-563
 *
 * Differences, all deliberate:
 *   - calls return b40c_status (0 = success) instead of cudaError_t;
 *   - VertexId and SizeT must be 32-bit (the reference's test_bfs uses int);
 *   - ExtractResults has a two-array overload returning depth and
 *     predecessor together; the one-array form keeps the reference meaning
 *     (predecessors when MARK_PREDECESSORS, else depths);
 *   - EnactorDirectionOptimizing and EnactorPersistent have no counterpart
 *     in the reference.
 */
#pragma once

#include <cstdint>
#include <type_traits>
#include <vector>

#include "../b40c_mi355x.h"

namespace b40c {
namespace graph {
namespace bfs {

enum FrontierType {
    VERTEX_FRONTIERS = B40C_VERTEX_FRONTIERS,
    EDGE_FRONTIERS = B40C_EDGE_FRONTIERS,
    MIXED_FRONTIERS = B40C_MIXED_FRONTIERS,
};

template <typename _VertexId = int, typename _SizeT = int, bool _MARK_PREDECESSORS = true>
class CsrProblem {
    static_assert(sizeof(_VertexId) == 4 && std::is_integral<_VertexId>::value, "VertexId must be a 32-bit integer");
    static_assert(sizeof(_SizeT) == 4 || sizeof(_SizeT) == 8, "SizeT must be a 32- or 64-bit integer");

public:
    typedef _VertexId VertexId;
    typedef _SizeT SizeT;
    static const bool MARK_PREDECESSORS = _MARK_PREDECESSORS;

    SizeT nodes = 0;
    SizeT edges = 0;

    CsrProblem() = default;
    CsrProblem(const CsrProblem&) = delete;
    CsrProblem& operator=(const CsrProblem&) = delete;
    ~CsrProblem() { b40c_problem_destroy(handle_); }

    // One GPU only: num_gpus != 1 returns B40C_ERR_UNSUPPORTED (the
    // multi-GPU partitioned path is not part of this library).
    // stream_from_host is accepted and ignored: the graph is copied to HBM.
    b40c_status FromHostProblem(bool stream_from_host, SizeT nodes_, SizeT edges_, VertexId* h_column_indices,
                                SizeT* h_row_offsets, int num_gpus = 1) {
        (void)stream_from_host;
        if (num_gpus != 1) return B40C_ERR_UNSUPPORTED;
        std::vector<int32_t> offs;
        const int32_t* p_offs = reinterpret_cast<const int32_t*>(h_row_offsets);
        if (sizeof(SizeT) != 4) {  // narrow 64-bit offsets; the device CSR is int32
            offs.resize((size_t)nodes_ + 1);
            for (size_t i = 0; i <= (size_t)nodes_; ++i) {
                if ((long long)h_row_offsets[i] > 0x7FFFFFFFll) return B40C_ERR_UNSUPPORTED;
                offs[i] = (int32_t)h_row_offsets[i];
            }
            p_offs = offs.data();
        }
        b40c_problem* p = nullptr;
        const b40c_status s = b40c_problem_create((int32_t)nodes_, (int64_t)edges_,
                                                  reinterpret_cast<const int32_t*>(h_column_indices), p_offs, 0, &p);
        if (s != B40C_OK) return s;
        b40c_problem_destroy(handle_);
        handle_ = p;
        nodes = nodes_;
        edges = edges_;
        return B40C_OK;
    }

    // Over a graph already resident on the device (no copy).
    b40c_status FromGraph(b40c_graph* g) {
        b40c_problem* p = nullptr;
        const b40c_status s = b40c_problem_create_from_graph(g, 0, &p);
        if (s != B40C_OK) return s;
        int32_t n = 0;
        int64_t m = 0;
        b40c_graph_info(g, &n, &m);
        b40c_problem_destroy(handle_);
        handle_ = p;
        nodes = (SizeT)n;
        edges = (SizeT)m;
        return B40C_OK;
    }

    b40c_status Reset(FrontierType frontier_type, double queue_sizing) {
        return b40c_problem_reset(handle_, (int32_t)frontier_type, queue_sizing);
    }

    // The reference's one-array form: predecessors when MARK_PREDECESSORS,
    // else depth labels.
    b40c_status ExtractResults(VertexId* h_labels) {
        int32_t* out = reinterpret_cast<int32_t*>(h_labels);
        return MARK_PREDECESSORS ? b40c_problem_extract(handle_, nullptr, out)
                                 : b40c_problem_extract(handle_, out, nullptr);
    }

    // Both arrays (either may be null).
    b40c_status ExtractResults(VertexId* h_depth, VertexId* h_pred) {
        return b40c_problem_extract(handle_, reinterpret_cast<int32_t*>(h_depth), reinterpret_cast<int32_t*>(h_pred));
    }

    b40c_problem* handle() const { return handle_; }

private:
    b40c_problem* handle_ = nullptr;
};

// Common body of the enactors; STRATEGY is a b40c_strategy value.
template <int STRATEGY, bool INSTRUMENT>
class EnactorBase {
public:
    explicit EnactorBase(bool DEBUG = false) : DEBUG(DEBUG) {
        status_ = b40c_enactor_create(STRATEGY, INSTRUMENT ? 1 : 0, &handle_);
    }
    EnactorBase(const EnactorBase&) = delete;
    EnactorBase& operator=(const EnactorBase&) = delete;
    ~EnactorBase() { b40c_enactor_destroy(handle_); }

    FrontierType GetFrontierType() const { return (FrontierType)b40c_enactor_frontier_type(handle_); }

    template <typename CsrProblem>
    b40c_status EnactSearch(CsrProblem& problem, typename CsrProblem::VertexId src, int max_grid_size = 0) {
        return enact(problem, src, max_grid_size, B40C_MODE_DEFAULT);
    }
    template <typename CsrProblem>
    b40c_status EnactFusedSearch(CsrProblem& problem, typename CsrProblem::VertexId src, int max_grid_size = 0) {
        return enact(problem, src, max_grid_size, B40C_MODE_FUSED);
    }
    template <typename CsrProblem>
    b40c_status EnactIterativeSearch(CsrProblem& problem, typename CsrProblem::VertexId src, int max_grid_size = 0) {
        return enact(problem, src, max_grid_size, B40C_MODE_ITERATIVE);
    }

    template <typename VertexId>
    b40c_status GetStatistics(long long& total_queued, VertexId& search_depth, double& avg_duty) {
        int64_t q = 0;
        int32_t d = 0;
        const b40c_status s = b40c_get_statistics(handle_, &q, &d, &avg_duty);
        total_queued = (long long)q;
        search_depth = (VertexId)d;
        return s;
    }

    // Runtime replacement for the reference's compile-time KernelPolicy.
    b40c_status SetTuning(const b40c_tuning& t) { return b40c_enactor_set_tuning(handle_, &t); }
    b40c_status GetTuning(b40c_tuning& t) const { return b40c_enactor_get_tuning(handle_, &t); }

    b40c_enactor* handle() const { return handle_; }
    b40c_status status() const { return status_; }

    bool DEBUG;

private:
    template <typename CsrProblem>
    b40c_status enact(CsrProblem& problem, typename CsrProblem::VertexId src, int max_grid_size, int32_t mode) {
        if (status_ != B40C_OK) return status_;
        return b40c_enact_search(handle_, problem.handle(), (int32_t)src, max_grid_size, mode);
    }

    b40c_enactor* handle_ = nullptr;
    b40c_status status_ = B40C_OK;
};

template <bool INSTRUMENT = false>
class EnactorContractExpand : public EnactorBase<B40C_CONTRACT_EXPAND, INSTRUMENT> {
public:
    explicit EnactorContractExpand(bool DEBUG = false) : EnactorBase<B40C_CONTRACT_EXPAND, INSTRUMENT>(DEBUG) {}
};

template <bool INSTRUMENT = false>
class EnactorExpandContract : public EnactorBase<B40C_EXPAND_CONTRACT, INSTRUMENT> {
public:
    explicit EnactorExpandContract(bool DEBUG = false) : EnactorBase<B40C_EXPAND_CONTRACT, INSTRUMENT>(DEBUG) {}
};

template <bool INSTRUMENT = false>
class EnactorTwoPhase : public EnactorBase<B40C_TWO_PHASE, INSTRUMENT> {
public:
    explicit EnactorTwoPhase(bool DEBUG = false) : EnactorBase<B40C_TWO_PHASE, INSTRUMENT>(DEBUG) {}
};

template <bool INSTRUMENT = false>
class EnactorHybrid : public EnactorBase<B40C_HYBRID, INSTRUMENT> {
public:
    explicit EnactorHybrid(bool DEBUG = false) : EnactorBase<B40C_HYBRID, INSTRUMENT>(DEBUG) {}
};

template <bool INSTRUMENT = false>
class EnactorPersistent : public EnactorBase<B40C_PERSISTENT, INSTRUMENT> {
public:
    explicit EnactorPersistent(bool DEBUG = false) : EnactorBase<B40C_PERSISTENT, INSTRUMENT>(DEBUG) {}
};

template <bool INSTRUMENT = false>
class EnactorDirectionOptimizing : public EnactorBase<B40C_DIRECTION_OPTIMIZING, INSTRUMENT> {
public:
    explicit EnactorDirectionOptimizing(bool DEBUG = false)
        : EnactorBase<B40C_DIRECTION_OPTIMIZING, INSTRUMENT>(DEBUG) {}
};

}  // namespace bfs
}  // namespace graph
}  // namespace b40c
