// bfs_internal.hpp -- types shared by the kernel translation unit and the
// C-ABI translation unit.  Not installed; the public interface is
// include/b40c_bfs.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b40c_hip {

struct Control;
struct HubEntry;

// Phases of one step of the level loop.
enum Phase : int {
    kPhaseExpandContract = 0,  // vertex frontier -> vertex frontier (gather + claim)
    kPhaseContractExpand = 1,  // edge frontier -> edge frontier (claim + gather)
    kPhaseContract = 2,        // edge frontier -> vertex frontier (claim)
    kPhaseExpand = 3,          // vertex frontier -> edge frontier (gather)
    kPhaseFilter = 4,          // edge frontier -> edge frontier (visited-bitmask cull)
    kPhaseBottomUp = 5         // frontier bitmap -> every unvisited vertex looks for a parent
};

struct GraphArgs {
    const int *row;   // [V + 1]
    const int *col;   // [E]
    int nodes;
    const int *rrow;  // reverse (in-neighbour) CSR for bottom-up levels;
    const int *rcol;  // equal to row / col when the graph is symmetric
};

struct SearchArgs {
    int *depth;            // [V]
    int *pred;             // [V] or null
    unsigned *visited;     // [ceil(V / 32)]
    int *q[2];             // ping-pong frontier queues
    int *pq[2];            // predecessor queues (null unless marking predecessors)
    unsigned cap;          // capacity of each queue, in elements
    Control *ctl;
    HubEntry *hubs[2];     // grid-tier lists, by step & 1
    unsigned hub_cap;      // entries per list
    unsigned *fbm[2];      // frontier bitmaps of the bottom-up levels, by iteration & 1
    unsigned char *marks;  // [32 * ceil(V / 32)] discovery marks of large top-down levels, clear between steps
    int nodes;
    unsigned long long total_edges;
    int track_edges;       // accumulate frontier edge counts (direction switching)
    int do_alpha, do_beta; // top-down -> bottom-up when m_f * alpha > m_u; back when n_f * beta < V
    int wave_t, cta_t, grid_t, fast_t, sat_t;
    // Hybrid two-phase launches: a contract step whose input is at most
    // sat_t hands the search back to the persistent contract-expand kernel
    // (sets Control::hybrid_stop), and every step launched after it returns.
    int hybrid_steps;
    int instrument;
    int grab_max;          // most tiles taken per tile-counter atomic
    int coherent_filter;   // claim pre-test reads the device-coherent visited word
    int spread;            // few-member top-down levels split their edges evenly over the grid
    int fast_visited;      // the fast path claims in an LDS copy of the visited bitmap (when it fits)
    // Chained compaction of contract and filter steps: one status word per
    // tile, tagged with epoch scan_epoch + step (never 0: cleared words read
    // as unpublished).  Null tile_status: wave-staged atomic reservation.
    unsigned long long *tile_status;
    unsigned scan_epoch;
    // Search start folded into a persistent launch (init_src >= 0): the
    // kernel clears the status head, queues init_src, labels it
    // (init_label) and, with init_reset, rewrites every label first --
    // what launch_init_search does as a launch of its own.
    int init_src;
    int init_reset;
    int init_label;
    // Pinned host copy of the control block's status head, written by a
    // persistent launch when its level loop ends (null: none).
    Control *host_head;
};

// Loop position: step index (selects queues and counters), BFS iteration
// (depth written by the next claim), and phase (two-phase cycle).
struct Cursor {
    int step;
    int iteration;
    int phase;
};

// Strategy of a persistent launch.
enum LoopKind : int { kLoopEC = 0, kLoopCE = 1, kLoopTP = 2, kLoopDO = 3 };

// Host launchers (bfs_kernels.hip).
// Clears the control block and queues src; with reset, also rewrites every
// label (depth, predecessor, visited bit) in the same launch.
hipError_t launch_init_search(const SearchArgs &s, const int *row, int src, int label_source, bool reset,
                              hipStream_t st);
// Sets the frontier bitmap fbm[iteration & 1] from the queue of step c.
hipError_t launch_build_frontier_bitmap(const SearchArgs &s, Cursor c, int grid, hipStream_t st, int *launches);
// Bitmap fbm[c.iteration & 1] -> input queue of step c (bottom-up -> top-down).
hipError_t launch_queue_from_bitmap(const SearchArgs &s, Cursor c, int grid, hipStream_t st, int *launches);
hipError_t launch_reset_labels(int *depth, int *pred, unsigned *visited, int nodes, hipStream_t st);
// One step as a kernel launch; gathering phases are followed by the
// grid-tier launch for rows published to the hub list.
hipError_t launch_step(int phase, bool pred, bool exact, const GraphArgs &g, const SearchArgs &s, Cursor c,
                       int grid, hipStream_t st, int *launches);
hipError_t launch_persistent(int loop, bool pred, bool exact, const GraphArgs &g, const SearchArgs &s, Cursor c,
                             int grid, hipStream_t st);
// Largest co-resident grid of the persistent kernel for `loop` (0 on error).
int persistent_max_blocks_per_cu(int loop, bool pred, bool exact);
int step_max_blocks_per_cu(bool pred);
int barrier_bench_max_blocks_per_cu();
hipError_t launch_barrier_bench(Control *ctl, int iterations, int grid, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
hipError_t launch_wave_selftest(const int *in, int *out_scan, int *out_rank, int n, hipStream_t st);

}  // namespace b40c_hip

#include "b40c_bfs.h"

namespace b40c_hip {
// Persistent grids stay well inside co-residency: at most half the reported
// occupancy and at most kMaxPersistentPerCu workgroups per CU (single-source,
// batched and barrier-benchmark launches alike).
constexpr int kMaxPersistentPerCu = 4;
// Grid of a persistent launch on num_cus CUs under that rule.
int persistent_grid_for(int num_cus, int loop, bool pred, bool exact, int blocks_per_cu, int max_grid_size);
// Launches a persistent kernel, one whose workgroups meet at grid barriers
// and so must all be resident at once.  A cooperative launch has the runtime
// guarantee that.  The device has one cooperative queue, held by the first
// process that makes a cooperative launch; in any other process the launch
// reports hipErrorIllegalState, and the kernel then goes out as a plain
// launch of the same grid -- which every caller sizes at most half the
// kernel's occupancy (and at most kMaxPersistentPerCu per CU), so it is
// resident at once unless other long-running work holds the CUs.  After the
// first such refusal the process launches plainly.
hipError_t launch_coresident(const void *kernel, dim3 grid, dim3 block, void **args, hipStream_t st);

// Error helpers (they record the message bfs_last_error_message returns).
bfs_status host_fail(bfs_status s, const char *msg);
bfs_status host_hip_fail(hipError_t e, const char *what);
// The tunables of a search (thresholds, switch points, grabs; no buffers)
// from a config, with the search start left to the caller (init_src = -1).
void apply_search_config(SearchArgs &s, const bfs_config &cfg, int grid);
}  // namespace b40c_hip
