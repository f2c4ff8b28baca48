// b40c_bfs.hpp -- header-only C++ layer over the C ABI (b40c_bfs.h) that
// re-creates the reference's class and method names, so a test/bfs-style
// driver ports by changing its includes and its error type:
//
//   b40c::graph::bfs::CsrProblem<VertexId, SizeT, MARK_PREDECESSORS>
//       <- b40c/graph/bfs/csr_problem.cuh:68-751
//   b40c::graph::bfs::Enactor{ExpandContract,ContractExpand,TwoPhase,Hybrid}<INSTRUMENT>
//       <- b40c/graph/bfs/enactor_{expand_contract,contract_expand,two_phase,hybrid}.cuh
//
// Methods return bfs_status (0 = success) where the reference returned
// cudaError_t.  Only int32 vertex ids and sizes, one GPU, and device-resident
// graphs are supported.
#pragma once

#include "b40c_bfs.h"

#include <cstdio>
#include <type_traits>

namespace b40c {
namespace graph {
namespace bfs {

enum FrontierType {
    VERTEX_FRONTIERS = BFS_VERTEX_FRONTIERS,
    EDGE_FRONTIERS = BFS_EDGE_FRONTIERS,
    MIXED_FRONTIERS = BFS_MIXED_FRONTIERS,
};

inline bfs_status Perror(bfs_status s, const char *what, const char *file, int line)
{
    if (s != BFS_OK) {
        std::fprintf(stderr, "[%s, %d] %s (bfs status %d: %s: %s)\n", file, line, what, (int) s,
                     bfs_status_string(s), bfs_last_error_message());
        std::fflush(stderr);
    }
    return s;
}

template <typename VertexId, typename SizeT, bool MARK_PREDECESSORS>
struct CsrProblem {
    static_assert(std::is_same<VertexId, int>::value && std::is_same<SizeT, int>::value,
                  "int32 vertex ids and sizes only");
    bfs_problem *handle = nullptr;
    SizeT nodes = 0, edges = 0;
    int device = 0;

    CsrProblem() = default;
    CsrProblem(const CsrProblem &) = delete;
    CsrProblem &operator=(const CsrProblem &) = delete;
    ~CsrProblem() { bfs_problem_destroy(handle); }

    // csr_problem.cuh:328-334.  stream_from_host (zero-copy host CSR) and
    // num_gpus > 1 are not supported.
    bfs_status FromHostProblem(bool stream_from_host, SizeT n, SizeT e, VertexId *h_column_indices,
                               SizeT *h_row_offsets, int num_gpus)
    {
        if (stream_from_host || num_gpus != 1) return BFS_ERR_UNSUPPORTED;
        bfs_problem_destroy(handle);
        handle = nullptr;
        nodes = n;
        edges = e;
        return Perror(bfs_problem_create(h_row_offsets, h_column_indices, n, e, MARK_PREDECESSORS ? 1 : 0,
                                         device, &handle),
                      "CsrProblem FromHostProblem failed", __FILE__, __LINE__);
    }

    // csr_problem.cuh:547-549
    bfs_status Reset(FrontierType ft, double queue_sizing)
    {
        return Perror(bfs_problem_reset(handle, (bfs_frontier_type) ft, queue_sizing), "CsrProblem Reset failed",
                      __FILE__, __LINE__);
    }

    // csr_problem.cuh:255
    bfs_status ExtractResults(VertexId *h_label)
    {
        return Perror(bfs_extract_labels(handle, h_label), "CsrProblem ExtractResults failed", __FILE__, __LINE__);
    }
};

class EnactorBase {
  protected:
    bfs_strategy strategy_;
    bool debug_;
    bool instrument_;
    bfs_stats stats_{};
    EnactorBase(bfs_strategy s, bool debug, bool instrument) : strategy_(s), debug_(debug), instrument_(instrument) {}

    // INSTRUMENT enactors (e.g. enactor_contract_expand.cuh:60-61) turn on the
    // problem's duty statistics (bfs_config::instrument) for their searches.
    template <typename Problem>
    bfs_status Enact(Problem &p, int src, int fused, int max_grid_size)
    {
        bfs_config cfg;
        bfs_status s = bfs_problem_get_config(p.handle, &cfg);
        if (s == BFS_OK && cfg.instrument != (instrument_ ? 1 : 0)) {
            cfg.instrument = instrument_ ? 1 : 0;
            s = bfs_problem_set_config(p.handle, &cfg);
        }
        if (s != BFS_OK) return Perror(s, "Enactor configuration failed", __FILE__, __LINE__);
        s = bfs_enact(p.handle, strategy_, fused, src, max_grid_size, &stats_);
        if (debug_ && s == BFS_OK)
            std::printf("DEBUG: %d launches, %.3f ms\n", stats_.kernel_launches, stats_.elapsed_ms);
        return Perror(s, "Enactor search failed", __FILE__, __LINE__);
    }

  public:
    // enactor_base.cuh:72
    FrontierType GetFrontierType() const { return (FrontierType) bfs_strategy_frontier_type(strategy_); }

    // e.g. enactor_contract_expand.cuh:236-250
    template <typename VertexId>
    void GetStatistics(long long &total_queued, VertexId &search_depth, double &avg_duty) const
    {
        total_queued = stats_.total_queued;
        search_depth = stats_.search_depth;
        avg_duty = stats_.avg_duty;
    }
    double ElapsedMillis() const { return stats_.elapsed_ms; }
};

template <bool INSTRUMENT>
class EnactorExpandContract : public EnactorBase {
  public:
    explicit EnactorExpandContract(bool debug = false) : EnactorBase(BFS_EXPAND_CONTRACT, debug, INSTRUMENT) {}
    template <typename Problem>
    bfs_status EnactFusedSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 1, max_grid_size);
    }
    template <typename Problem>
    bfs_status EnactIterativeSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 0, max_grid_size);
    }
};

template <bool INSTRUMENT>
class EnactorContractExpand : public EnactorBase {
  public:
    explicit EnactorContractExpand(bool debug = false) : EnactorBase(BFS_CONTRACT_EXPAND, debug, INSTRUMENT) {}
    template <typename Problem>
    bfs_status EnactFusedSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 1, max_grid_size);
    }
    template <typename Problem>
    bfs_status EnactIterativeSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 0, max_grid_size);
    }
};

template <bool INSTRUMENT>
class EnactorTwoPhase : public EnactorBase {
  public:
    explicit EnactorTwoPhase(bool debug = false) : EnactorBase(BFS_TWO_PHASE, debug, INSTRUMENT) {}
    template <typename Problem>
    bfs_status EnactFusedSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 1, max_grid_size);
    }
    template <typename Problem>
    bfs_status EnactIterativeSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 0, max_grid_size);
    }
};

// Addition (no reference counterpart): top-down expand-contract levels mixed
// with bottom-up levels (bfs_config do_alpha / do_beta switch points).
template <bool INSTRUMENT>
class EnactorDirectionOptimizing : public EnactorBase {
  public:
    explicit EnactorDirectionOptimizing(bool debug = false) : EnactorBase(BFS_DIRECTION_OPTIMIZING, debug, INSTRUMENT) {}
    template <typename Problem>
    bfs_status EnactFusedSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 1, max_grid_size);
    }
    template <typename Problem>
    bfs_status EnactIterativeSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 0, max_grid_size);
    }
};

template <bool INSTRUMENT>
class EnactorHybrid : public EnactorBase {
  public:
    explicit EnactorHybrid(bool debug = false) : EnactorBase(BFS_HYBRID, debug, INSTRUMENT) {}
    template <typename Problem>
    bfs_status EnactSearch(Problem &p, int src, int max_grid_size = 0)
    {
        return Enact(p, src, 1, max_grid_size);
    }
};

}  // namespace bfs
}  // namespace graph
}  // namespace b40c
