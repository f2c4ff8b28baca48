// bfs_batch.hpp -- device-side types of batched BFS (many single-source
// queries over a pool of resident graphs, advanced level by level together).
// Shared by bfs_batch.hip (kernels and the batch / server host code).
#pragma once

#include "bfs_device.hpp"

namespace b40c_hip {

constexpr int kMaxSlots = 256;  // queries advanced by one launch

struct BatchGraph {
    const int *row, *col;    // CSR
    const int *rrow, *rcol;  // in-neighbour CSR (aliases when symmetric)
    int nodes;
    int pad;
    unsigned long long edges;
};

// One query's state.  Counters rotate by the query's own step exactly as in
// the single-source Control block.
struct alignas(128) BatchSlot {
    int graph, src, active, done;
    int step, iteration, phase, overflow;
    int next_phase, ran;      // direction chosen for the level in flight, and whether it had work
    int *depth;
    unsigned *visited;
    unsigned *fbm[2];         // frontier bitmaps of bottom-up levels, by iteration & 1
    int *q[2];
    unsigned cap;             // queue capacity (elements)
    unsigned qlen[3];
    unsigned long long fedges[3];
    unsigned long long visited_edges;
    unsigned long long t_begin, t_end;  // wall clock at seeding / completion
    unsigned ovf[3];          // per-step overflow of the output queue, by step % 3
    int bits;                 // fbm[iteration & 1] holds exactly the current frontier
    unsigned long long vbase; // sum of fedges over the steps before the current one
};

struct alignas(256) BatchControl {
    unsigned steal[2];        // tile counters by level parity
    int level;                // levels completed by the batch
    int pad0;
    unsigned long long t_start;
    unsigned bar_pad[26];
    unsigned long long hub_alloc[2];  // grid tier: (entries << 32) | chunks, by level & 1
    unsigned hub_steal[2];            // grid tier: chunk counter, by level & 1
    int hub_chunk[2];                 // grid tier: edges per chunk, by level & 1
    unsigned pad1[26];
    // grid barrier state (same layout discipline as Control)
    Control bar;
    // per-level record: wall clock at the level start, top-down and
    // bottom-up tiles planned
    unsigned long long level_clock[kLevelLog];
    unsigned level_td_tiles[kLevelLog];
    unsigned level_bu_tiles[kLevelLog];
};

struct BatchArgs {
    const BatchGraph *graphs;
    BatchSlot *slots;
    BatchControl *ctl;
    int nslots;
    int wave_t, cta_t;
    int direction;            // 0: top-down only, 1: direction-optimizing
    int do_alpha, do_beta;
    HubEntry *hubs[2];        // grid-tier rows of all queries, by level & 1
    unsigned hub_cap;
    int grid_t;               // rows at least this long go to the grid tier
    int grab_max;             // most tiles taken per tile-counter atomic
    int bu_tile;              // kBlock-vertex blocks per bottom-up tile
    int strided;              // grab layout: tiles gridDim apart, else consecutive
    int coherent_filter;      // claim pre-test reads the device-coherent visited word
};

}  // namespace b40c_hip
