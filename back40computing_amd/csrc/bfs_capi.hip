// bfs_capi.hip -- C ABI of libb40c_hip.so (declared in include/b40c_bfs.h).
//
// Host side of the single-GPU BFS: problem storage (the reference's
// CsrProblem, b40c/graph/bfs/csr_problem.cuh:68-751) and the four enactors
// (b40c/graph/bfs/enactor_*.cuh) as strategy drivers over the kernels in
// bfs_kernels.hip.
#include "b40c_bfs.h"
#include "bfs_device.hpp"
#include "bfs_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "b40c_graph.h"

using namespace b40c_hip;

namespace {

thread_local std::string g_last_error;

bfs_status fail(bfs_status s, const std::string &msg)
{
    g_last_error = msg;
    return s;
}

bfs_status hip_fail(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return BFS_ERR_OOM;
    if (e == hipErrorCooperativeLaunchTooLarge || e == hipErrorLaunchOutOfResources) return BFS_ERR_LAUNCH;
    return BFS_ERR_HIP;
}

#define HIP_TRY(expr, what)                         \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// Defaults: the autotuner's pick for direction-optimizing search on R-MAT
// scale 22, edge factor 16 (back40computing_amd/autotune.py; its table is
// back40computing_amd/tuned.json).
const bfs_config kDefaults = {
    /*wave_gather_threshold*/ 64,
    /*cta_gather_threshold*/ 512,
    /*fast_path_threshold*/ 1024,
    /*saturation_threshold*/ 0,   // 0: persistent grid size * 256
    /*blocks_per_cu*/ 1,
    /*cull_mode*/ BFS_CULL_EXACT,
    /*instrument*/ 0,
    /*grid_gather_threshold*/ 8192,
    /*do_alpha*/ 8,
    /*do_beta*/ 48,
    /*tile_grab_max*/ 32,
    /*bu_tile_blocks*/ 1,
    /*grab_strided*/ 1,
    /*claim_filter*/ 1,
    /*spread_steps*/ 1,
    /*compaction*/ 1,
    /*fast_visited*/ 1};

constexpr unsigned kHubCapacity = 1u << 16;  // grid-tier entries per step list

}  // namespace

struct bfs_problem {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int32_t nodes = 0;
    int64_t edges = 0;
    bool mark_pred = false;
    int num_cus = 0;

    int *d_row = nullptr;
    int *d_col = nullptr;
    int *d_rrow = nullptr;       // reverse CSR (aliases d_row / d_col when symmetric)
    int *d_rcol = nullptr;
    bool symmetric = true;
    int *d_depth = nullptr;
    int *d_pred = nullptr;
    unsigned *d_visited = nullptr;
    int *d_q[2] = {nullptr, nullptr};
    int *d_pq[2] = {nullptr, nullptr};
    size_t q_cap = 0;            // allocated elements per queue
    // chained compaction: one status word per kBlock-item tile of a queue,
    // and the epoch base of the next search (advanced by a stride that
    // exceeds any search's step count; the words are cleared on wrap)
    unsigned long long *d_tile_status = nullptr;
    unsigned scan_epoch = 1;
    unsigned cap = 0;            // capacity in use for the current frontier type
    int frontier_type = -1;      // set by reset
    bool dirty = false;          // labels written since the last reset

    Control *d_ctl = nullptr;
    Control *h_ctl = nullptr;    // pinned mirror
    Control *h_ctl_dev = nullptr;  // its device-side address
    HubEntry *d_hubs = nullptr;  // two grid-tier lists of kHubCapacity entries
    unsigned *d_fbm = nullptr;   // two frontier bitmaps (bottom-up levels)
    unsigned char *d_marks = nullptr;  // discovery marks of large top-down levels (32 per visited word)
    int wall_khz = 100000;
    bfs_config cfg = kDefaults;
};

namespace {

// Persistent grids follow kMaxPersistentPerCu (bfs_internal.hpp).  A launch
// at full occupancy (8 per CU) went silent once and was never shown to be
// safe, so that shape is not issued.
int persistent_grid(const bfs_problem *p, int loop, int max_grid_size)
{
    return persistent_grid_for(p->num_cus, loop, p->mark_pred, p->cfg.cull_mode == BFS_CULL_EXACT,
                               p->cfg.blocks_per_cu, max_grid_size);
}

int step_grid(const bfs_problem *p, int max_grid_size)
{
    int occ = step_max_blocks_per_cu(p->mark_pred);
    int grid = p->num_cus * std::max(1, occ);
    if (max_grid_size > 0) grid = std::min(grid, max_grid_size);
    return std::max(grid, 1);
}

SearchArgs search_args(const bfs_problem *p, int grid)
{
    SearchArgs s;
    apply_search_config(s, p->cfg, grid);
    s.depth = p->d_depth;
    s.pred = p->d_pred;
    s.visited = p->d_visited;
    s.q[0] = p->d_q[0];
    s.q[1] = p->d_q[1];
    s.pq[0] = p->d_pq[0];
    s.pq[1] = p->d_pq[1];
    s.cap = p->cap;
    s.ctl = p->d_ctl;
    s.hubs[0] = p->d_hubs;
    s.hubs[1] = p->d_hubs + kHubCapacity;
    s.hub_cap = kHubCapacity;
    size_t words = ((size_t) p->nodes + 31) / 32;
    s.fbm[0] = p->d_fbm;
    s.fbm[1] = p->d_fbm + words;
    s.marks = p->d_marks;
    s.tile_status = p->cfg.compaction == 2 ? p->d_tile_status : nullptr;
    s.scan_epoch = p->scan_epoch;
    s.nodes = p->nodes;
    s.total_edges = (unsigned long long) p->edges;
    return s;
}

bfs_status read_control(bfs_problem *p)
{
    HIP_TRY(hipMemcpyAsync(p->h_ctl, p->d_ctl, kControlHead, hipMemcpyDeviceToHost, p->stream), "read control block");
    HIP_TRY(hipStreamSynchronize(p->stream), "synchronize");
    return BFS_OK;
}

// Launches per-step kernels following `next` from cursor c until the
// frontier is empty.  Steps are issued in chunks; an empty frontier makes
// the remaining launches of a chunk return immediately.
template <typename Next>
bfs_status run_iterative(bfs_problem *p, const GraphArgs &g, const SearchArgs &s, Cursor &c, int grid,
                         int *launches, Next next, bool (*stop)(const Control &, Cursor, unsigned))
{
    const bool exact = p->cfg.cull_mode == BFS_CULL_EXACT;
    const int chunk = 6;
    while (true) {
        for (int k = 0; k < chunk; k++) {
            HIP_TRY(launch_step(c.phase, p->mark_pred, exact, g, s, c, grid, p->stream, launches),
                    "launch step kernel");
            c = next(c);
        }
        bfs_status st = read_control(p);
        if (st != BFS_OK) return st;
        const Control &h = *p->h_ctl;
        unsigned n = h.qlen[c.step % 3];
        if (h.overflow || n == 0) return BFS_OK;
        if (stop && stop(h, c, n)) return BFS_OK;
    }
}

Cursor next_same(Cursor c)
{
    c.step++;
    c.iteration++;
    return c;
}

Cursor next_two_phase(Cursor c)
{
    c.step++;
    if (c.phase == kPhaseContract) c.phase = kPhaseExpand;
    else if (c.phase == kPhaseExpand) c.phase = kPhaseFilter;
    else { c.phase = kPhaseContract; c.iteration++; }
    return c;
}

bfs_status enact_hybrid(bfs_problem *p, const GraphArgs &g, const SearchArgs &s, int pgrid, int sgrid,
                        int *launches)
{
    const bool exact = p->cfg.cull_mode == BFS_CULL_EXACT;
    // Two-phase levels per host check: the launches of a chunk past the
    // hand-back (or past an empty frontier) return at once, so the host
    // waits once per chunk rather than once per level
    // (enactor_hybrid.cuh:483-492 runs ahead of its done flag the same way).
    const int chunk = 4;
    SearchArgs ts = s;
    ts.hybrid_steps = 1;
    Cursor c{0, 0, kPhaseContractExpand};
    while (true) {
        // Persistent contract-expand until the edge frontier saturates
        // (enactor_hybrid.cuh:335-384).
        HIP_TRY(launch_persistent(kLoopCE, p->mark_pred, exact, g, s, c, pgrid, p->stream),
                "launch persistent contract-expand");
        (*launches)++;
        bfs_status st = read_control(p);
        if (st != BFS_OK) return st;
        const Control &h = *p->h_ctl;
        if (h.exit_reason != kExitSaturated) return BFS_OK;
        // Two-phase levels: filter -> contract -> expand, until the filtered
        // edge frontier fits the persistent kernel again
        // (enactor_hybrid.cuh:386-532); the contract step that finds it so
        // records where and stops the rest of the chunk.
        c = Cursor{h.exit_step, h.exit_iteration, kPhaseFilter};
        HIP_TRY(hipMemsetAsync(&p->d_ctl->hybrid_stop, 0, sizeof(unsigned), p->stream), "clear hand-back flag");
        while (true) {
            for (int k = 0; k < chunk; k++) {
                HIP_TRY(launch_step(kPhaseFilter, p->mark_pred, exact, g, ts, c, sgrid, p->stream, launches),
                        "launch filter");
                Cursor cc{c.step + 1, c.iteration, kPhaseContract};
                HIP_TRY(launch_step(kPhaseContract, p->mark_pred, exact, g, ts, cc, sgrid, p->stream, launches),
                        "launch contract");
                Cursor ce{c.step + 2, c.iteration, kPhaseExpand};
                HIP_TRY(launch_step(kPhaseExpand, p->mark_pred, exact, g, ts, ce, sgrid, p->stream, launches),
                        "launch expand");
                c = Cursor{c.step + 3, c.iteration + 1, kPhaseFilter};
            }
            st = read_control(p);
            if (st != BFS_OK) return st;
            if (p->h_ctl->overflow) return BFS_OK;
            if (p->h_ctl->hybrid_stop) {
                c = Cursor{p->h_ctl->exit_step, p->h_ctl->exit_iteration, kPhaseContractExpand};
                break;
            }
            if (p->h_ctl->qlen[c.step % 3] == 0) return BFS_OK;
        }
    }
}

// Per-level launches of the direction-optimizing search: the host reads
// the frontier counters after every level and applies the same switch rule
// as the persistent kernel.
bfs_status enact_direction_iterative(bfs_problem *p, const GraphArgs &g, const SearchArgs &s, int grid,
                                     int *launches)
{
    const bool exact = p->cfg.cull_mode == BFS_CULL_EXACT;
    Cursor c{0, 0, kPhaseExpandContract};
    unsigned long long prev_mf = 0;  // the previous frontier's edges
    while (true) {
        bfs_status st = read_control(p);
        if (st != BFS_OK) return st;
        const Control &h = *p->h_ctl;
        unsigned n = h.qlen[c.step % 3];
        if (h.overflow || n == 0) return BFS_OK;
        unsigned long long mf = h.fedges[c.step % 3];
        unsigned long long mu = s.total_edges > h.visited_edges ? s.total_edges - h.visited_edges : 0ull;
        // the persistent loop's rule (bottom-up only while the frontier grows)
        int dir = c.phase;
        if (c.phase == kPhaseExpandContract) {
            if (mf > prev_mf && mf * (unsigned long long) s.do_alpha > mu) dir = kPhaseBottomUp;
        } else if ((unsigned long long) n * (unsigned long long) s.do_beta < (unsigned long long) s.nodes) {
            dir = kPhaseExpandContract;
        }
        prev_mf = mf;
        if (dir == kPhaseBottomUp && c.phase != kPhaseBottomUp)
            HIP_TRY(launch_build_frontier_bitmap(s, c, grid, p->stream, launches), "build frontier bitmap");
        else if (dir != kPhaseBottomUp && c.phase == kPhaseBottomUp)
            HIP_TRY(launch_queue_from_bitmap(s, c, grid, p->stream, launches), "queue from bitmap");
        c.phase = dir;
        HIP_TRY(launch_step(c.phase, p->mark_pred, exact, g, s, c, grid, p->stream, launches), "launch step");
        c.step++;
        c.iteration++;
    }
}

// Whether the edge multiset equals its reverse, by two independent 64-bit
// commutative hashes of (u, v) and (v, u) over all edges.  Unequal sums prove
// asymmetry; equal sums are taken as symmetry (a false positive needs a
// 128-bit collision).
bool csr_is_symmetric(const int32_t *row, const int32_t *col, int32_t nodes)
{
    auto mix = [](uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    int threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<uint64_t> acc(4 * threads, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t] {
            uint64_t a = 0, b = 0, c = 0, d = 0;
            for (int32_t u = t; u < nodes; u += threads) {
                for (int32_t e = row[u]; e < row[u + 1]; e++) {
                    uint64_t v = (uint32_t) col[e];
                    uint64_t fw = ((uint64_t) (uint32_t) u << 32) | v, bw = (v << 32) | (uint32_t) u;
                    a += mix(fw);
                    b += mix(bw);
                    c += mix(fw ^ 0x5851F42D4C957F2Dull);
                    d += mix(bw ^ 0x5851F42D4C957F2Dull);
                }
            }
            acc[4 * t] = a;
            acc[4 * t + 1] = b;
            acc[4 * t + 2] = c;
            acc[4 * t + 3] = d;
        });
    }
    for (auto &th : pool) th.join();
    uint64_t a = 0, b = 0, c = 0, d = 0;
    for (int t = 0; t < threads; t++) {
        a += acc[4 * t];
        b += acc[4 * t + 1];
        c += acc[4 * t + 2];
        d += acc[4 * t + 3];
    }
    return a == b && c == d;
}

bfs_status validate_csr(const int32_t *row, const int32_t *col, int32_t nodes, int64_t edges)
{
    if (!row || (edges > 0 && !col) || nodes <= 0 || edges < 0 || edges > INT32_MAX)
        return fail(BFS_ERR_INVALID_ARG, "invalid CSR arguments");
    if (row[0] != 0 || (int64_t) row[nodes] != edges) return fail(BFS_ERR_INVALID_ARG, "row offsets do not span [0, edges]");
    for (int32_t v = 0; v < nodes; v++)
        if (row[v + 1] < row[v]) return fail(BFS_ERR_INVALID_ARG, "row offsets are not monotone");
    for (int64_t e = 0; e < edges; e++)
        if (col[e] < 0 || col[e] >= nodes) return fail(BFS_ERR_INVALID_ARG, "column index out of range");
    return BFS_OK;
}

void free_queues(bfs_problem *p)
{
    for (int i = 0; i < 2; i++) {
        if (p->d_q[i]) (void) hipFree(p->d_q[i]);
        if (p->d_pq[i]) (void) hipFree(p->d_pq[i]);
        p->d_q[i] = p->d_pq[i] = nullptr;
    }
    if (p->d_tile_status) (void) hipFree(p->d_tile_status);
    p->d_tile_status = nullptr;
    p->q_cap = 0;
}

}  // namespace

extern "C" {

const char *bfs_status_string(bfs_status s)
{
    switch (s) {
    case BFS_OK: return "success";
    case BFS_ERR_INVALID_ARG: return "invalid argument";
    case BFS_ERR_OOM: return "out of device memory";
    case BFS_ERR_OVERFLOW: return "Frontier queue overflow. Please increase queue-sizing factor.";
    case BFS_ERR_UNSUPPORTED: return "unsupported";
    case BFS_ERR_LAUNCH: return "kernel launch failed (grid too large for co-residency?)";
    case BFS_ERR_HIP: return "HIP runtime error";
    case BFS_ERR_NO_DEVICE: return "no HIP device";
    case BFS_ERR_REJECTED: return "request rejected by admission control";
    case BFS_ERR_TIMEOUT: return "timed out";
    case BFS_ERR_NOT_READY: return "not ready";
    }
    return "unknown status";
}

const char *bfs_last_error_message(void) { return g_last_error.c_str(); }

bfs_frontier_type bfs_strategy_frontier_type(bfs_strategy s)
{
    return (s == BFS_EXPAND_CONTRACT || s == BFS_DIRECTION_OPTIMIZING) ? BFS_VERTEX_FRONTIERS : BFS_EDGE_FRONTIERS;
}

bfs_status bfs_config_default(bfs_config *cfg)
{
    if (!cfg) return fail(BFS_ERR_INVALID_ARG, "null config");
    *cfg = kDefaults;
    return BFS_OK;
}

bfs_status bfs_device_count(int32_t *count)
{
    if (!count) return fail(BFS_ERR_INVALID_ARG, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return BFS_OK;
}

bfs_status bfs_device_synchronize(int32_t device)
{
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return BFS_OK;
}

bfs_status bfs_problem_create(const int32_t *h_row_offsets, const int32_t *h_column_indices, int32_t nodes,
                              int64_t edges, int32_t mark_predecessors, int32_t device, bfs_problem **out)
{
    if (!out) return fail(BFS_ERR_INVALID_ARG, "null output handle");
    *out = nullptr;
    bfs_status vs = validate_csr(h_row_offsets, h_column_indices, nodes, edges);
    if (vs != BFS_OK) return vs;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(BFS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(BFS_ERR_INVALID_ARG, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");

    bfs_problem *p = new bfs_problem();
    p->device = device;
    p->nodes = nodes;
    p->edges = edges;
    p->mark_pred = mark_predecessors != 0;
    auto cleanup = [&](bfs_status s) {
        bfs_problem_destroy(p);
        return s;
    };
    hipError_t e;
    if ((e = hipDeviceGetAttribute(&p->num_cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return cleanup(hip_fail(e, "query CU count"));
    if (hipDeviceGetAttribute(&p->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        p->wall_khz <= 0)
        p->wall_khz = 100000;
    if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess)
        return cleanup(hip_fail(e, "create stream"));
    if ((e = hipEventCreate(&p->ev_start)) != hipSuccess || (e = hipEventCreate(&p->ev_stop)) != hipSuccess)
        return cleanup(hip_fail(e, "create events"));
    size_t words = ((size_t) nodes + 31) / 32;
    if ((e = hipMalloc(&p->d_row, sizeof(int) * ((size_t) nodes + 1))) != hipSuccess ||
        (e = hipMalloc(&p->d_col, sizeof(int) * (size_t) std::max<int64_t>(edges, 1))) != hipSuccess ||
        (e = hipMalloc(&p->d_depth, sizeof(int) * (size_t) nodes)) != hipSuccess ||
        (e = hipMalloc(&p->d_visited, sizeof(unsigned) * words)) != hipSuccess ||
        (e = hipMalloc(&p->d_ctl, sizeof(Control))) != hipSuccess ||
        (e = hipMalloc(&p->d_hubs, sizeof(HubEntry) * 2 * kHubCapacity)) != hipSuccess ||
        (e = hipMalloc(&p->d_fbm, sizeof(unsigned) * 2 * words)) != hipSuccess ||
        (e = hipMalloc(&p->d_marks, 32 * words)) != hipSuccess)
        return cleanup(hip_fail(e, "allocate problem"));
    // marks are clear between steps (each compaction clears what it reads)
    if ((e = hipMemset(p->d_marks, 0, 32 * words)) != hipSuccess) return cleanup(hip_fail(e, "clear marks"));
    if (p->mark_pred && (e = hipMalloc(&p->d_pred, sizeof(int) * (size_t) nodes)) != hipSuccess)
        return cleanup(hip_fail(e, "allocate predecessors"));
    if ((e = hipHostMalloc(&p->h_ctl, sizeof(Control), hipHostMallocDefault)) != hipSuccess)
        return cleanup(hip_fail(e, "allocate pinned control block"));
    if ((e = hipHostGetDevicePointer(reinterpret_cast<void **>(&p->h_ctl_dev), p->h_ctl, 0)) != hipSuccess)
        return cleanup(hip_fail(e, "map pinned control block"));
    // the grid barrier's counters must start at zero (searches leave them so)
    if ((e = hipMemset(p->d_ctl, 0, sizeof(Control))) != hipSuccess)
        return cleanup(hip_fail(e, "clear control block"));
    if ((e = hipMemcpy(p->d_row, h_row_offsets, sizeof(int) * ((size_t) nodes + 1), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(hip_fail(e, "copy row offsets"));
    if (edges > 0 &&
        (e = hipMemcpy(p->d_col, h_column_indices, sizeof(int) * (size_t) edges, hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(hip_fail(e, "copy column indices"));

    // Bottom-up levels scan in-neighbours: reuse the CSR when the graph is
    // symmetric, otherwise upload its transpose.
    p->symmetric = csr_is_symmetric(h_row_offsets, h_column_indices, nodes);
    if (p->symmetric) {
        p->d_rrow = p->d_row;
        p->d_rcol = p->d_col;
    } else {
        std::vector<int32_t> src((size_t) edges);
        for (int32_t v = 0; v < nodes; v++)
            for (int32_t k = h_row_offsets[v]; k < h_row_offsets[v + 1]; k++) src[k] = v;
        b40c_csr t{};
        if (b40c_graph_from_coo(nodes, edges, h_column_indices, src.data(), 0, 0, 0, &t) != 0)
            return cleanup(fail(BFS_ERR_OOM, "transpose construction failed"));
        bool ok = hipMalloc(&p->d_rrow, sizeof(int) * ((size_t) nodes + 1)) == hipSuccess &&
                  hipMalloc(&p->d_rcol, sizeof(int) * (size_t) std::max<int64_t>(edges, 1)) == hipSuccess &&
                  hipMemcpy(p->d_rrow, t.row_offsets, sizeof(int) * ((size_t) nodes + 1), hipMemcpyHostToDevice) ==
                      hipSuccess &&
                  (edges == 0 || hipMemcpy(p->d_rcol, t.column_indices, sizeof(int) * (size_t) edges,
                                           hipMemcpyHostToDevice) == hipSuccess);
        b40c_graph_free(&t);
        if (!ok) return cleanup(fail(BFS_ERR_OOM, "allocate or copy the transpose"));
    }
    *out = p;
    return BFS_OK;
}

bfs_status bfs_problem_reset(bfs_problem *p, bfs_frontier_type ft, double queue_sizing)
{
    if (!p) return fail(BFS_ERR_INVALID_ARG, "null problem");
    if (ft < BFS_VERTEX_FRONTIERS || ft > BFS_MIXED_FRONTIERS || !(queue_sizing > 0.0))
        return fail(BFS_ERR_INVALID_ARG, "invalid frontier type or queue sizing");
    HIP_TRY(hipSetDevice(p->device), "hipSetDevice");
    double base = ft == BFS_VERTEX_FRONTIERS ? (double) p->nodes : (double) std::max<int64_t>(p->edges, p->nodes);
    double want = std::ceil(queue_sizing * base);
    if (want < 1.0) want = 1.0;
    if (want > 4.0e9) return fail(BFS_ERR_INVALID_ARG, "queue capacity exceeds 32-bit counters");
    size_t cap = (size_t) want;
    if (cap > p->q_cap) {
        free_queues(p);
        for (int i = 0; i < 2; i++) {
            HIP_TRY(hipMalloc(&p->d_q[i], sizeof(int) * cap), "allocate frontier queue");
            if (p->mark_pred) HIP_TRY(hipMalloc(&p->d_pq[i], sizeof(int) * cap), "allocate predecessor queue");
        }
        const size_t words = cap / kBlock + 2;
        HIP_TRY(hipMalloc(&p->d_tile_status, sizeof(unsigned long long) * words), "allocate tile status words");
        HIP_TRY(hipMemsetAsync(p->d_tile_status, 0, sizeof(unsigned long long) * words, p->stream),
                "clear tile status words");
        p->scan_epoch = 1;
        p->q_cap = cap;
    }
    p->cap = (unsigned) cap;
    p->frontier_type = ft;
    HIP_TRY(launch_reset_labels(p->d_depth, p->d_pred, p->d_visited, p->nodes, p->stream), "reset labels");
    HIP_TRY(hipStreamSynchronize(p->stream), "synchronize");
    p->dirty = false;
    return BFS_OK;
}

bfs_status bfs_problem_set_config(bfs_problem *p, const bfs_config *cfg)
{
    if (!p || !cfg) return fail(BFS_ERR_INVALID_ARG, "null argument");
    bfs_config c = *cfg;
    if (c.wave_gather_threshold == 0) c.wave_gather_threshold = kDefaults.wave_gather_threshold;
    if (c.cta_gather_threshold == 0) c.cta_gather_threshold = kDefaults.cta_gather_threshold;
    if (c.fast_path_threshold == 0) c.fast_path_threshold = kDefaults.fast_path_threshold;
    if (c.blocks_per_cu == 0) c.blocks_per_cu = kDefaults.blocks_per_cu;
    if (c.grid_gather_threshold == 0) c.grid_gather_threshold = kDefaults.grid_gather_threshold;
    if (c.grid_gather_threshold < 0) c.grid_gather_threshold = -1;  // grid tier off
    if (c.claim_filter == 0) c.claim_filter = kDefaults.claim_filter;
    if (c.spread_steps == 0) c.spread_steps = kDefaults.spread_steps;
    if (c.compaction == 0) c.compaction = kDefaults.compaction;
    if (c.compaction < 1 || c.compaction > 2) return fail(BFS_ERR_INVALID_ARG, "unknown compaction mode");
    if (c.fast_visited == 0) c.fast_visited = kDefaults.fast_visited;
    if (c.fast_visited < 1 || c.fast_visited > 2) return fail(BFS_ERR_INVALID_ARG, "unknown fast_visited mode");
    if (c.do_alpha == 0) c.do_alpha = kDefaults.do_alpha;
    if (c.do_beta == 0) c.do_beta = kDefaults.do_beta;
    if (c.do_alpha < 1 || c.do_beta < 1) return fail(BFS_ERR_INVALID_ARG, "invalid direction switch points");
    if (c.wave_gather_threshold < 1 || c.cta_gather_threshold < 1 || c.blocks_per_cu < 1 ||
        (c.cull_mode != BFS_CULL_EXACT && c.cull_mode != BFS_CULL_BEST_EFFORT))
        return fail(BFS_ERR_INVALID_ARG, "invalid config");
    if (c.fast_path_threshold < -1) c.fast_path_threshold = -1;
    p->cfg = c;
    return BFS_OK;
}

bfs_status bfs_problem_get_config(const bfs_problem *p, bfs_config *cfg)
{
    if (!p || !cfg) return fail(BFS_ERR_INVALID_ARG, "null argument");
    *cfg = p->cfg;
    return BFS_OK;
}

bfs_status bfs_enact(bfs_problem *p, bfs_strategy strategy, int32_t fused, int32_t src, int32_t max_grid_size,
                     bfs_stats *stats)
{
    if (!p) return fail(BFS_ERR_INVALID_ARG, "null problem");
    if (strategy < BFS_EXPAND_CONTRACT || strategy > BFS_DIRECTION_OPTIMIZING)
        return fail(BFS_ERR_INVALID_ARG, "unknown strategy");
    if (src < 0 || src >= p->nodes) return fail(BFS_ERR_INVALID_ARG, "source vertex out of range");
    if (p->frontier_type < 0) return fail(BFS_ERR_INVALID_ARG, "problem not reset");
    bfs_frontier_type need = bfs_strategy_frontier_type(strategy);
    if (p->frontier_type != need && p->frontier_type != BFS_MIXED_FRONTIERS)
        return fail(BFS_ERR_INVALID_ARG, "problem was reset for another frontier type");
    HIP_TRY(hipSetDevice(p->device), "hipSetDevice");
    const bool reset = p->dirty;
    p->dirty = true;

    const bool exact = p->cfg.cull_mode == BFS_CULL_EXACT;
    int loop = strategy == BFS_EXPAND_CONTRACT ? kLoopEC
               : strategy == BFS_TWO_PHASE       ? kLoopTP
               : strategy == BFS_DIRECTION_OPTIMIZING ? kLoopDO
                                                 : kLoopCE;
    const bool vertex_frontier = strategy == BFS_EXPAND_CONTRACT || strategy == BFS_DIRECTION_OPTIMIZING;
    int pgrid = persistent_grid(p, loop, max_grid_size);
    if (pgrid > p->num_cus * std::max(1, persistent_max_blocks_per_cu(loop, p->mark_pred, exact)))
        return fail(BFS_ERR_LAUNCH, "grid exceeds co-resident capacity");
    int sgrid = step_grid(p, max_grid_size);
    GraphArgs g{p->d_row, p->d_col, p->nodes, p->d_rrow, p->d_rcol};
    // Chained compaction tags each step's tile words with scan_epoch + step.
    // A search has fewer than 3 (V + 2) steps (every level but the last
    // labels a vertex; two-phase levels take three steps).
    const unsigned long long stride = 3ull * ((unsigned long long) p->nodes + 2);
    if (p->cfg.compaction == 2 && (unsigned long long) p->scan_epoch + 2 * stride >= (1ull << 30)) {
        const size_t words = p->q_cap / kBlock + 2;
        HIP_TRY(hipMemsetAsync(p->d_tile_status, 0, sizeof(unsigned long long) * words, p->stream),
                "clear tile status words");
        p->scan_epoch = 1;
    }
    SearchArgs s = search_args(p, pgrid);
    if (p->cfg.compaction == 2) p->scan_epoch += (unsigned) stride;
    if (strategy != BFS_HYBRID) s.sat_t = -1;

    int launches = 0;
    if (!vertex_frontier) s.track_edges = 0;
    // One persistent launch per search starts the search itself and leaves
    // the status head in pinned memory; the other paths start it with a
    // launch of its own and copy the head afterwards.
    const bool folded = fused && strategy != BFS_HYBRID;
    if (folded) {
        s.init_src = src;
        s.init_reset = reset ? 1 : 0;
        s.init_label = vertex_frontier ? 1 : 0;
        // instrumented cycle sums arrive after the kernel's copy of the head
        if (!s.instrument) s.host_head = p->h_ctl_dev;
    } else {
        HIP_TRY(launch_init_search(s, p->d_row, src, vertex_frontier ? 1 : 0, reset, p->stream), "init search");
    }
    HIP_TRY(hipEventRecord(p->ev_start, p->stream), "record start");
    bfs_status st = BFS_OK;
    Cursor c{0, 0, kPhaseExpandContract};
    switch (strategy) {
    case BFS_EXPAND_CONTRACT:
        c.phase = kPhaseExpandContract;
        if (fused) {
            HIP_TRY(launch_persistent(kLoopEC, p->mark_pred, exact, g, s, c, pgrid, p->stream), "launch persistent");
            launches++;
        } else {
            st = run_iterative(p, g, s, c, sgrid, &launches, next_same, nullptr);
        }
        break;
    case BFS_CONTRACT_EXPAND:
        c.phase = kPhaseContractExpand;
        if (fused) {
            HIP_TRY(launch_persistent(kLoopCE, p->mark_pred, exact, g, s, c, pgrid, p->stream), "launch persistent");
            launches++;
        } else {
            st = run_iterative(p, g, s, c, sgrid, &launches, next_same, nullptr);
        }
        break;
    case BFS_TWO_PHASE:
        c.phase = kPhaseContract;
        if (fused) {
            HIP_TRY(launch_persistent(kLoopTP, p->mark_pred, exact, g, s, c, pgrid, p->stream), "launch persistent");
            launches++;
        } else {
            st = run_iterative(p, g, s, c, sgrid, &launches, next_two_phase, nullptr);
        }
        break;
    case BFS_HYBRID:
        st = enact_hybrid(p, g, s, pgrid, sgrid, &launches);
        break;
    case BFS_DIRECTION_OPTIMIZING:
        c.phase = kPhaseExpandContract;
        if (fused) {
            HIP_TRY(launch_persistent(kLoopDO, p->mark_pred, exact, g, s, c, pgrid, p->stream), "launch persistent");
            launches++;
        } else {
            st = enact_direction_iterative(p, g, s, sgrid, &launches);
        }
        break;
    }
    if (st != BFS_OK) return st;
    HIP_TRY(hipEventRecord(p->ev_stop, p->stream), "record stop");
    // the status head of the control block, queued behind the search: one
    // wait covers both
    if (!s.host_head)
        HIP_TRY(hipMemcpyAsync(p->h_ctl, p->d_ctl, kControlHead, hipMemcpyDeviceToHost, p->stream),
                "read control block");
    HIP_TRY(hipStreamSynchronize(p->stream), "synchronize");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p->ev_start, p->ev_stop), "elapsed time");
    const Control &h = *p->h_ctl;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->total_queued = (int64_t) h.total_queued;
        stats->search_depth = h.max_depth + 1;
        stats->avg_duty = h.life_cycles ? (double) h.busy_cycles / (double) h.life_cycles : 0.0;
        stats->elapsed_ms = ms;
        stats->kernel_launches = launches;
    }
    if (h.overflow) return fail(BFS_ERR_OVERFLOW, "Frontier queue overflow. Please increase queue-sizing factor.");
    return BFS_OK;
}

bfs_status bfs_extract_labels(const bfs_problem *p, int32_t *h_labels)
{
    if (!p || !h_labels) return fail(BFS_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(p->device), "hipSetDevice");
    const int *src = p->mark_pred ? p->d_pred : p->d_depth;
    HIP_TRY(hipMemcpy(h_labels, src, sizeof(int) * (size_t) p->nodes, hipMemcpyDeviceToHost), "copy labels");
    return BFS_OK;
}

bfs_status bfs_extract_depths(const bfs_problem *p, int32_t *h_depths)
{
    if (!p || !h_depths) return fail(BFS_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(p->device), "hipSetDevice");
    HIP_TRY(hipMemcpy(h_depths, p->d_depth, sizeof(int) * (size_t) p->nodes, hipMemcpyDeviceToHost), "copy depths");
    return BFS_OK;
}

void bfs_problem_destroy(bfs_problem *p)
{
    if (!p) return;
    (void) hipSetDevice(p->device);
    if (p->stream) (void) hipStreamSynchronize(p->stream);
    free_queues(p);
    if (p->d_rrow && p->d_rrow != p->d_row) (void) hipFree(p->d_rrow);
    if (p->d_rcol && p->d_rcol != p->d_col) (void) hipFree(p->d_rcol);
    int *bufs[] = {p->d_row, p->d_col, p->d_depth, p->d_pred};
    for (int *b : bufs)
        if (b) (void) hipFree(b);
    if (p->d_visited) (void) hipFree(p->d_visited);
    if (p->d_ctl) (void) hipFree(p->d_ctl);
    if (p->d_hubs) (void) hipFree(p->d_hubs);
    if (p->d_fbm) (void) hipFree(p->d_fbm);
    if (p->d_marks) (void) hipFree(p->d_marks);
    if (p->h_ctl) (void) hipHostFree(p->h_ctl);
    if (p->ev_start) (void) hipEventDestroy(p->ev_start);
    if (p->ev_stop) (void) hipEventDestroy(p->ev_stop);
    if (p->stream) (void) hipStreamDestroy(p->stream);
    delete p;
}

bfs_status bfs_problem_device_info(const bfs_problem *p, int32_t *num_cus, int32_t *grid_size, int64_t *device_bytes)
{
    if (!p) return fail(BFS_ERR_INVALID_ARG, "null problem");
    if (num_cus) *num_cus = p->num_cus;
    if (grid_size) *grid_size = persistent_grid(p, kLoopEC, 0);
    if (device_bytes) {
        int64_t b = 4 * ((int64_t) p->nodes + 1) + 4 * p->edges + 4 * (int64_t) p->nodes * (p->mark_pred ? 2 : 1) +
                    (4 + 8 + 32) * (((int64_t) p->nodes + 31) / 32) + (int64_t) sizeof(Control) +
                    (int64_t) p->q_cap * 4 * 2 * (p->mark_pred ? 2 : 1);
        *device_bytes = b;
    }
    return BFS_OK;
}

bfs_status bfs_problem_device_arrays(const bfs_problem *p, const int32_t **d_row_offsets,
                                    const int32_t **d_column_indices, const int32_t **d_in_row_offsets,
                                    const int32_t **d_in_column_indices, int32_t **d_depths)
{
    if (!p) return fail(BFS_ERR_INVALID_ARG, "null problem");
    if (d_row_offsets) *d_row_offsets = p->d_row;
    if (d_column_indices) *d_column_indices = p->d_col;
    if (d_in_row_offsets) *d_in_row_offsets = p->d_rrow;
    if (d_in_column_indices) *d_in_column_indices = p->d_rcol;
    if (d_depths) *d_depths = p->d_depth;
    return BFS_OK;
}

bfs_status bfs_get_level_log(const bfs_problem *p, int32_t max_steps, uint32_t *lengths, double *start_us,
                             int32_t *steps)
{
    if (!p || !steps || max_steps < 0) return fail(BFS_ERR_INVALID_ARG, "invalid argument");
    HIP_TRY(hipSetDevice(p->device), "hipSetDevice");
    HIP_TRY(hipMemcpy(p->h_ctl, p->d_ctl, sizeof(Control), hipMemcpyDeviceToHost), "read control block");
    const Control &h = *p->h_ctl;
    int k = 0;
    for (; k < kLevelLog && k < max_steps; k++) {
        if (k > 0 && h.level_clock[k] == 0) break;
        if (lengths) lengths[k] = h.level_len[k];
        if (start_us) start_us[k] = 1000.0 * (double) (h.level_clock[k] - h.level_clock[0]) / p->wall_khz;
    }
    *steps = k;
    return BFS_OK;
}

bfs_status bfs_barrier_benchmark(int32_t device, int32_t blocks_per_cu, int32_t iterations, double *us_per_barrier,
                                 int32_t *grid_size)
{
    if (!us_per_barrier || iterations < 1 || blocks_per_cu < 1 || blocks_per_cu > 8)
        return fail(BFS_ERR_INVALID_ARG, "invalid argument");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(BFS_ERR_NO_DEVICE, "no HIP device");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "query CU count");
    // the co-residency rule of every persistent launch (persistent_grid)
    const int occ = barrier_bench_max_blocks_per_cu();
    if (blocks_per_cu > std::max(1, std::min(kMaxPersistentPerCu, occ / 2)))
        return fail(BFS_ERR_INVALID_ARG, "blocks_per_cu above the persistent-grid cap min(4, occupancy / 2)");
    int grid = cus * blocks_per_cu;
    if (grid_size) *grid_size = grid;
    Control *d_ctl = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bfs_status rc = BFS_OK;
    do {
        hipError_t e;
        if ((e = hipMalloc(&d_ctl, sizeof(Control))) != hipSuccess) { rc = hip_fail(e, "allocate"); break; }
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) { rc = hip_fail(e, "stream"); break; }
        if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
            rc = hip_fail(e, "events");
            break;
        }
        if ((e = hipMemsetAsync(d_ctl, 0, sizeof(Control), st)) != hipSuccess) { rc = hip_fail(e, "clear"); break; }
        // Warm-up launch, then the timed launch.
        if ((e = launch_barrier_bench(d_ctl, 1, grid, st)) != hipSuccess) { rc = hip_fail(e, "launch"); break; }
        if ((e = hipEventRecord(e0, st)) != hipSuccess) { rc = hip_fail(e, "record"); break; }
        if ((e = launch_barrier_bench(d_ctl, iterations, grid, st)) != hipSuccess) { rc = hip_fail(e, "launch"); break; }
        if ((e = hipEventRecord(e1, st)) != hipSuccess) { rc = hip_fail(e, "record"); break; }
        if ((e = hipEventSynchronize(e1)) != hipSuccess) { rc = hip_fail(e, "synchronize"); break; }
        float ms = 0;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) { rc = hip_fail(e, "elapsed"); break; }
        *us_per_barrier = 1000.0 * ms / iterations;
    } while (0);
    if (d_ctl) (void) hipFree(d_ctl);
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    if (st) (void) hipStreamDestroy(st);
    return rc;
}

bfs_status bfs_copy_calibration(int32_t device, int64_t bytes, double *ms)
{
    if (!ms || bytes <= 0 || bytes % 16) return fail(BFS_ERR_INVALID_ARG, "invalid argument");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(BFS_ERR_NO_DEVICE, "no HIP device");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bfs_status rc = BFS_OK;
    do {
        hipError_t e;
        if ((e = hipMalloc(&a, bytes)) != hipSuccess || (e = hipMalloc(&b, bytes)) != hipSuccess) {
            rc = hip_fail(e, "allocate");
            break;
        }
        if ((e = hipMemset(a, 1, bytes)) != hipSuccess) { rc = hip_fail(e, "fill"); break; }
        if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
            rc = hip_fail(e, "events");
            break;
        }
        if ((e = hipEventRecord(e0, nullptr)) != hipSuccess) { rc = hip_fail(e, "record"); break; }
        if ((e = launch_copy(a, b, (size_t) bytes, nullptr)) != hipSuccess) { rc = hip_fail(e, "launch"); break; }
        if ((e = hipEventRecord(e1, nullptr)) != hipSuccess) { rc = hip_fail(e, "record"); break; }
        if ((e = hipEventSynchronize(e1)) != hipSuccess) { rc = hip_fail(e, "synchronize"); break; }
        float t = 0;
        if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) { rc = hip_fail(e, "elapsed"); break; }
        *ms = t;
    } while (0);
    if (a) (void) hipFree(a);
    if (b) (void) hipFree(b);
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    return rc;
}

bfs_status bfs_selftest_wave_primitives(int32_t device, int32_t *mismatches)
{
    if (!mismatches) return fail(BFS_ERR_INVALID_ARG, "null argument");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(BFS_ERR_NO_DEVICE, "no HIP device");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    const int n = 64 * 37 + 5;
    std::vector<int> in(n), scan(n), rank(n);
    uint32_t x = 12345u;
    for (int i = 0; i < n; i++) {
        x = x * 1664525u + 1013904223u;
        in[i] = (int) ((x >> 16) % 97);
    }
    int *d_in = nullptr, *d_scan = nullptr, *d_rank = nullptr;
    HIP_TRY(hipMalloc(&d_in, sizeof(int) * n), "allocate");
    HIP_TRY(hipMalloc(&d_scan, sizeof(int) * n), "allocate");
    HIP_TRY(hipMalloc(&d_rank, sizeof(int) * n), "allocate");
    HIP_TRY(hipMemcpy(d_in, in.data(), sizeof(int) * n, hipMemcpyHostToDevice), "copy");
    HIP_TRY(launch_wave_selftest(d_in, d_scan, d_rank, n, nullptr), "launch");
    HIP_TRY(hipDeviceSynchronize(), "synchronize");
    HIP_TRY(hipMemcpy(scan.data(), d_scan, sizeof(int) * n, hipMemcpyDeviceToHost), "copy");
    HIP_TRY(hipMemcpy(rank.data(), d_rank, sizeof(int) * n, hipMemcpyDeviceToHost), "copy");
    (void) hipFree(d_in);
    (void) hipFree(d_scan);
    (void) hipFree(d_rank);
    int bad = 0;
    for (int w = 0; w * 64 < n; w++) {
        int acc = 0, odd = 0;
        for (int l = 0; l < 64 && w * 64 + l < n; l++) {
            int i = w * 64 + l;
            acc += in[i];
            if (scan[i] != acc || rank[i] != odd) bad++;
            odd += in[i] & 1;
        }
    }
    *mismatches = bad;
    return bad == 0 ? BFS_OK : fail(BFS_ERR_HIP, "wave primitive mismatch");
}

}  // extern "C"

// Helpers shared with the batch / server translation unit.
namespace b40c_hip {

bfs_status host_fail(bfs_status s, const char *msg) { return fail(s, msg); }

void apply_search_config(SearchArgs &s, const bfs_config &cfg, int grid)
{
    s.wave_t = cfg.wave_gather_threshold;
    s.cta_t = cfg.cta_gather_threshold;
    s.grid_t = cfg.grid_gather_threshold > 0 ? cfg.grid_gather_threshold : 0x7fffffff;
    s.grab_max = cfg.tile_grab_max > 0 ? cfg.tile_grab_max : 1;
    s.coherent_filter = cfg.claim_filter == 2;
    s.spread = cfg.spread_steps == 1;
    s.fast_visited = cfg.fast_visited == 1;
    s.fast_t = cfg.fast_path_threshold;
    // the reference leaves at grid * 512 items for 128-thread CTAs
    // (enactor_hybrid.cuh:571): four items per thread of the grid
    s.sat_t = cfg.saturation_threshold > 0 ? cfg.saturation_threshold
              : (cfg.saturation_threshold < 0 ? -1 : grid * kBlock * 4);
    s.hybrid_steps = 0;
    s.instrument = cfg.instrument;
    s.tile_status = nullptr;
    s.scan_epoch = 0;
    s.track_edges = 1;
    s.do_alpha = cfg.do_alpha;
    s.do_beta = cfg.do_beta;
    s.init_src = -1;
    s.init_reset = 0;
    s.init_label = 0;
    s.host_head = nullptr;
}

int persistent_grid_for(int num_cus, int loop, bool pred, bool exact, int blocks_per_cu, int max_grid_size)
{
    int occ = persistent_max_blocks_per_cu(loop, pred, exact);
    int cap = std::max(1, std::min(kMaxPersistentPerCu, occ / 2));
    int per_cu = std::min(cap, std::max(1, blocks_per_cu));
    int grid = num_cus * per_cu;
    if (max_grid_size > 0) grid = std::min(grid, max_grid_size);
    return std::max(grid, 1);
}

bfs_status host_hip_fail(hipError_t e, const char *what) { return hip_fail(e, what); }

bfs_status host_validate_csr(const int32_t *row, const int32_t *col, int32_t nodes, int64_t edges)
{
    return validate_csr(row, col, nodes, edges);
}

bool host_csr_is_symmetric(const int32_t *row, const int32_t *col, int32_t nodes)
{
    return csr_is_symmetric(row, col, nodes);
}

}  // namespace b40c_hip
