/*
 * b40c_bfs.h -- C ABI of the MI355X (gfx950) breadth-first-search library
 * libb40c_hip.so.  Plain C types only; device memory is owned by the library.
 *
 * The reference exposes this path as header-only C++ templates that its
 * driver instantiates (test/bfs/test_bfs.cu:600-738).  Each entry point below
 * replaces one of those calls:
 *
 *   bfs_problem_create        <- CsrProblem::FromHostProblem
 *                                (b40c/graph/bfs/csr_problem.cuh:328-540)
 *   bfs_problem_reset         <- CsrProblem::Reset (csr_problem.cuh:547-750)
 *   bfs_extract_labels        <- CsrProblem::ExtractResults (csr_problem.cuh:255-273)
 *   bfs_problem_destroy       <- CsrProblem::~CsrProblem / GraphSlice::~GraphSlice
 *                                (csr_problem.cuh:152-172)
 *   bfs_strategy_frontier_type<- EnactorBase::GetFrontierType (b40c/graph/bfs/enactor_base.cuh:72)
 *   bfs_enact                 <- EnactorContractExpand::EnactFusedSearch
 *                                (b40c/graph/bfs/enactor_contract_expand.cuh:338-407),
 *                                EnactorExpandContract::EnactFusedSearch
 *                                (enactor_expand_contract.cuh:335-373),
 *                                EnactorTwoPhase::EnactIterativeSearch / EnactFusedSearch
 *                                (enactor_two_phase.cuh:360, 623-765),
 *                                EnactorHybrid::EnactSearch (enactor_hybrid.cuh:559-751),
 *                                followed by Enactor*::GetStatistics
 *                                (e.g. enactor_contract_expand.cuh:236-250)
 *   bfs_status                <- cudaError_t return codes of the above, with the
 *                                overflow case of enactor_contract_expand.cuh:318-324
 *                                as BFS_ERR_OVERFLOW
 *   bfs_config                <- the compile-time KernelPolicy parameters
 *                                (b40c/graph/bfs/contract_expand_atomic/kernel_policy.cuh:61-279)
 *                                as runtime tunables
 *
 * Additions with no reference counterpart (marked "addition" below):
 * depth extraction in predecessor mode, the autotuner hooks, batched BFS over
 * a resident graph pool, and the continuous-batching server.
 *
 * Threading: one problem handle is used by one host thread at a time; calls
 * are synchronous on the handle's own HIP stream (the reference ran on the
 * legacy default stream and returned after completion).
 */
#ifndef B40C_BFS_H
#define B40C_BFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bfs_status {
    BFS_OK = 0,
    BFS_ERR_INVALID_ARG = 1,
    BFS_ERR_OOM = 2,
    BFS_ERR_OVERFLOW = 3,      /* frontier queue overflow: raise queue_sizing */
    BFS_ERR_UNSUPPORTED = 4,
    BFS_ERR_LAUNCH = 5,        /* e.g. persistent grid larger than co-resident capacity */
    BFS_ERR_HIP = 6,
    BFS_ERR_NO_DEVICE = 7,
    BFS_ERR_REJECTED = 8,      /* serving: admission refused (HBM budget) */
    BFS_ERR_TIMEOUT = 9,       /* serving: wait timed out */
    BFS_ERR_NOT_READY = 10     /* serving: result not finished yet */
} bfs_status;

/* Same numbering as the reference driver's Strategy enum (test/bfs/test_bfs.cu:72-79). */
typedef enum bfs_strategy {
    BFS_EXPAND_CONTRACT = 0,
    BFS_CONTRACT_EXPAND = 1,
    BFS_TWO_PHASE = 2,
    BFS_HYBRID = 3,
    /* addition: expand-contract top-down levels mixed with bottom-up levels,
       switching direction by frontier edge and vertex counts (Beamer et al.,
       SC'12); switch points are the do_alpha / do_beta tunables */
    BFS_DIRECTION_OPTIMIZING = 4
} bfs_strategy;

/* Same numbering as FrontierType (b40c/graph/bfs/csr_problem.cuh:54-58). */
typedef enum bfs_frontier_type {
    BFS_VERTEX_FRONTIERS = 0,
    BFS_EDGE_FRONTIERS = 1,
    BFS_MIXED_FRONTIERS = 2
} bfs_frontier_type;

/* Visited-culling discipline. */
typedef enum bfs_cull_mode {
    BFS_CULL_EXACT = 0,        /* atomic visited bitmask: every vertex enqueued once */
    BFS_CULL_BEST_EFFORT = 1   /* reference discipline: racy bitmask + label check;
                                  duplicates possible, labels still exact */
} bfs_cull_mode;

/*
 * Runtime tunables (the reference's KernelPolicy template arguments).
 * Zero in any field selects the library default.
 */
typedef struct bfs_config {
    int32_t wave_gather_threshold;  /* row length >= this: one wave gathers the row   */
    int32_t cta_gather_threshold;   /* row length >= this: the whole workgroup gathers */
    int32_t fast_path_threshold;    /* frontier <= this: one workgroup runs the level
                                       with no grid barrier (-1 disables)             */
    int32_t saturation_threshold;   /* hybrid: edge frontier > this leaves the
                                       persistent contract-expand kernel (-1: never;
                                       default: 4 items per thread of the grid)       */
    int32_t blocks_per_cu;          /* persistent grid = CUs * this (capped by occupancy) */
    int32_t cull_mode;              /* bfs_cull_mode */
    int32_t instrument;             /* 1: collect per-workgroup duty statistics */
    int32_t grid_gather_threshold;  /* row length >= this: the row is split into
                                       2048-edge chunks shared by the whole grid */
    int32_t do_alpha;               /* direction-optimizing: go bottom-up when
                                       frontier edges * alpha > unexplored edges */
    int32_t do_beta;                /* direction-optimizing: back to top-down when
                                       frontier vertices * beta < V */
    int32_t tile_grab_max;          /* most tiles (one item per thread of a workgroup:
                                       512 single-source, 256 batched) a workgroup
                                       takes per tile-counter atomic; the grab is
                                       tiles / (grid * 8) clamped to [1, this] */
    int32_t bu_tile_blocks;         /* batched: most 256-vertex blocks per bottom-up
                                       tile (levels with few blocks use fewer) */
    int32_t grab_strided;           /* batched: 1 = a grab's tiles are gridDim apart
                                       (narrow in-flight window), 2 = consecutive */
    int32_t claim_filter;           /* exact claims: the visited-word test before the
                                       atomic reads 1 = the cached copy, 2 = the
                                       device-coherent value */
    int32_t spread_steps;           /* top-down levels of at most 1024 vertices with
                                       too many edges for one workgroup: 1 = every
                                       workgroup walks an equal share of their edges
                                       (one phase), 2 = tile and grid-tier phases */
    int32_t compaction;             /* output of contract and filter steps (two-phase,
                                       hybrid): 1 = each wave stages it in LDS and
                                       reserves queue space with one atomic per flush;
                                       2 = a single-pass chained scan over the step's
                                       tiles (decoupled look-back), which keeps the
                                       input order */
    int32_t fast_visited;           /* visited claims of the single-workgroup fast
                                       path: 1 = in a copy of the visited bitmap in
                                       the workgroup's LDS when the graph has at most
                                       BFS_FAST_VISITED_MAX vertices (else as 2),
                                       2 = atomics on the bitmap in device memory */
} bfs_config;

/* Most vertices for which fast_visited = 1 keeps the visited bitmap in LDS. */
#define BFS_FAST_VISITED_MAX 1103872

typedef struct bfs_stats {
    int64_t total_queued;    /* elements passed through the frontier queues */
    int32_t search_depth;    /* number of BFS levels (max depth + 1) */
    double avg_duty;         /* instrumented only: busy / (max busy * grid) */
    double elapsed_ms;       /* hipEvent time of the traversal (test_bfs.cu:691-693 scope) */
    int32_t kernel_launches; /* launches issued by this enact */
    int32_t reserved[3];
} bfs_stats;

typedef struct bfs_problem bfs_problem;

const char *bfs_status_string(bfs_status s);
/* addition: text of the last error raised on the calling thread. */
const char *bfs_last_error_message(void);
bfs_frontier_type bfs_strategy_frontier_type(bfs_strategy s);
bfs_status bfs_config_default(bfs_config *cfg);
bfs_status bfs_device_count(int32_t *count);
/* Synchronises the calling thread with all work on `device`. */
bfs_status bfs_device_synchronize(int32_t device);

/* Copies the host CSR (row_offsets[nodes+1], column_indices[edges]) to the device. */
bfs_status bfs_problem_create(const int32_t *h_row_offsets, const int32_t *h_column_indices,
                              int32_t nodes, int64_t edges, int32_t mark_predecessors,
                              int32_t device, bfs_problem **out);
/* Allocates frontier queues for `ft` (queue_sizing * V or * E elements) and
 * resets labels to -1 and the visited mask to 0. */
bfs_status bfs_problem_reset(bfs_problem *p, bfs_frontier_type ft, double queue_sizing);
bfs_status bfs_problem_set_config(bfs_problem *p, const bfs_config *cfg);
bfs_status bfs_problem_get_config(const bfs_problem *p, bfs_config *cfg);
/* Runs one search from `src`.  fused != 0 selects the persistent (single
 * launch, in-kernel grid barrier) form; fused == 0 the per-level launches.
 * Synchronous.  max_grid_size > 0 caps the grid. */
bfs_status bfs_enact(bfs_problem *p, bfs_strategy strategy, int32_t fused, int32_t src,
                     int32_t max_grid_size, bfs_stats *stats);
/* Depth labels, or predecessors (-2 source, -1 unreached) when the problem
 * was created with mark_predecessors. */
bfs_status bfs_extract_labels(const bfs_problem *p, int32_t *h_labels);
/* addition: depth labels in either mode. */
bfs_status bfs_extract_depths(const bfs_problem *p, int32_t *h_depths);
void bfs_problem_destroy(bfs_problem *p);

/* addition: device buffer addresses of a problem (for callers that keep their
 * own device data, and for profiling). */
bfs_status bfs_problem_device_info(const bfs_problem *p, int32_t *num_cus, int32_t *grid_size,
                                   int64_t *device_bytes);

/* addition: device addresses of a problem's CSR (as uploaded), its
 * in-neighbour CSR (the same arrays when the graph is symmetric) and its
 * depth labels.  Any pointer argument may be NULL. */
bfs_status bfs_problem_device_arrays(const bfs_problem *p, const int32_t **d_row_offsets,
                                    const int32_t **d_column_indices, const int32_t **d_in_row_offsets,
                                    const int32_t **d_in_column_indices, int32_t **d_depths);

/* addition: per-step record of the last enact (first 64 steps): input
 * frontier length and start time in microseconds relative to step 0.
 * Returns the number of steps recorded in *steps. */
bfs_status bfs_get_level_log(const bfs_problem *p, int32_t max_steps, uint32_t *lengths, double *start_us,
                             int32_t *steps);

/* addition: microbenchmark of the in-kernel grid barrier (the reference's
 * test/gbarrier/test_gbarrier.cu:43-116): `iterations` back-to-back barriers
 * in one persistent launch; returns microseconds per barrier. */
bfs_status bfs_barrier_benchmark(int32_t device, int32_t blocks_per_cu, int32_t iterations,
                                 double *us_per_barrier, int32_t *grid_size);

/* addition: streaming copy of `bytes` (multiple of 16) between two device
 * buffers by a plain kernel (b40c_calibration_copy_kernel), for scaling the
 * FETCH_SIZE / WRITE_SIZE hardware counters to bytes.  *ms: kernel time. */
bfs_status bfs_copy_calibration(int32_t device, int64_t bytes, double *ms);

/* addition: self-test of the wave-level prefix scan and ballot compaction
 * primitives against a scalar computation on the device.  Returns BFS_OK when
 * every lane agrees; *mismatches receives the number of disagreeing lanes. */
bfs_status bfs_selftest_wave_primitives(int32_t device, int32_t *mismatches);

#ifdef __cplusplus
}
#endif

#endif /* B40C_BFS_H */
