/*
 * b40c_bfs_batch.h -- batched BFS and the continuous-batching BFS server
 * (C ABI, exported by libb40c_hip.so).
 *
 * These are additions with no counterpart in the reference, which runs one
 * search at a time (test/bfs/test_bfs.cu:584-783).  Each query has the
 * single-source semantics of bfs_enact with the expand-contract or the
 * direction-optimizing strategy: exact depth labels (-1 unreached).
 *
 *   graph pool  - device-resident CSR graphs shared by batches and servers
 *   batch       - up to 256 queries over pool graphs advanced level by level
 *                 together: one persistent launch (mode 0), or per-level
 *                 launches captured into a hipGraph and replayed (mode 1)
 *   server      - a host thread that admits submitted queries into free
 *                 slots at level boundaries while earlier queries are still
 *                 running (continuous batching), bounded by an HBM budget for
 *                 per-query frontier and label memory
 */
#ifndef B40C_BFS_BATCH_H
#define B40C_BFS_BATCH_H

#include <stdint.h>

#include "b40c_bfs.h"
#include "b40c_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfs_graph_pool bfs_graph_pool;
typedef struct bfs_batch bfs_batch;
typedef struct bfs_server bfs_server;

typedef struct bfs_batch_stats {
    double elapsed_ms;       /* hipEvent time of the level loop (seeding excluded) */
    int32_t queries;
    int32_t levels;          /* batch levels executed */
    int32_t launches;
    int32_t reserved0;
    double latency_p50_ms;   /* per-query completion time from the loop start */
    double latency_p99_ms;
    double latency_max_ms;
} bfs_batch_stats;

bfs_status bfs_pool_create(int32_t device, bfs_graph_pool **out);
/* Copies a CSR to the device; returns its id (0, 1, 2, ...). */
bfs_status bfs_pool_add_graph(bfs_graph_pool *pool, const int32_t *h_row_offsets, const int32_t *h_column_indices,
                              int32_t nodes, int64_t edges, int32_t *graph_id);
/* Adds a CSR already in device memory (b40c_device_csr_* builders), copied
 * device to device; the in-neighbour CSR of an asymmetric graph is built on
 * the device too.  The caller keeps (and frees) its device graph. */
bfs_status bfs_pool_add_device_graph(bfs_graph_pool *pool, const b40c_device_csr *graph, int32_t *graph_id);
bfs_status bfs_pool_graph_info(const bfs_graph_pool *pool, int32_t graph_id, int32_t *nodes, int64_t *edges);
bfs_status bfs_pool_device_bytes(const bfs_graph_pool *pool, int64_t *bytes);
void bfs_pool_destroy(bfs_graph_pool *pool);

/* Bytes of per-query device state (labels, visited mask, two frontier
 * queues) for a graph of `nodes` vertices: the unit of server admission. */
int64_t bfs_query_state_bytes(int32_t nodes);

bfs_status bfs_batch_create(bfs_graph_pool *pool, int32_t max_queries, bfs_batch **out);
/* Library defaults for batches and servers (bfs_config_default with the
 * batched tile shape and switch points).  Zero fields passed to
 * bfs_batch_set_config take these values. */
bfs_status bfs_batch_config_default(bfs_config *cfg);
bfs_status bfs_batch_set_config(bfs_batch *b, const bfs_config *cfg);
/* Runs nq (<= max_queries, <= 256) queries to completion.  mode 0: one
 * persistent launch; mode 1: hipGraph-captured per-level launches. */
bfs_status bfs_batch_enact(bfs_batch *b, const int32_t *graph_ids, const int32_t *sources, int32_t nq,
                           int32_t direction_optimizing, int32_t mode, bfs_batch_stats *stats);
/* Completion time of every query of the last enact, ms from the loop start. */
bfs_status bfs_batch_query_latency(const bfs_batch *b, double *ms);
bfs_status bfs_batch_extract_depths(const bfs_batch *b, int32_t query, int32_t *h_depths);
/* Per-level record of the last enact (first 64 levels): top-down and
 * bottom-up tiles planned and the level start in microseconds. */
bfs_status bfs_batch_level_log(const bfs_batch *b, int32_t max_levels, uint32_t *td_tiles, uint32_t *bu_tiles,
                               double *start_us, int32_t *levels);
void bfs_batch_destroy(bfs_batch *b);

typedef struct bfs_server_config {
    int32_t max_slots;            /* queries resident on the device at once (<= 256) */
    int32_t levels_per_tick;      /* levels per replay of the captured level graph */
    int64_t hbm_budget_bytes;     /* device memory for per-query state */
    int32_t direction_optimizing; /* 1: direction-optimizing queries, 0: top-down */
    int32_t reserved[5];
} bfs_server_config;

typedef struct bfs_request_info {
    int64_t ticket;
    int32_t graph_id;
    int32_t src;
    int32_t levels;          /* BFS levels of the query (max depth + 1) */
    int32_t status;          /* bfs_status of the query */
    double submit_ms;        /* host times, ms since the server started */
    double admit_ms;
    double finish_ms;
} bfs_request_info;

typedef struct bfs_server_stats {
    int64_t submitted;
    int64_t admitted;
    int64_t completed;
    int64_t rejected;
    int64_t ticks;            /* level-graph replays */
    int64_t max_resident;     /* most queries resident at once */
    double busy_ms;           /* host time with at least one query resident */
    double tick_ms;           /* host time in level-graph replays (GPU busy) */
    double copy_ms;           /* result-copy thread time (overlaps the ticks) */
} bfs_server_stats;

bfs_status bfs_server_create(bfs_graph_pool *pool, const bfs_server_config *cfg, bfs_server **out);
/* Queues a query.  h_depths_out (may be NULL) receives the depth labels on
 * completion and must stay valid until bfs_server_wait returns for the
 * ticket.  BFS_ERR_REJECTED if the query's state can never fit the budget.
 * Thread-safe. */
bfs_status bfs_server_submit(bfs_server *s, int32_t graph_id, int32_t src, int32_t *h_depths_out, int64_t *ticket);
/* Waits for a ticket (timeout_ms < 0: forever).  BFS_ERR_TIMEOUT on timeout. */
bfs_status bfs_server_wait(bfs_server *s, int64_t ticket, int32_t timeout_ms, bfs_request_info *info);
bfs_status bfs_server_get_stats(bfs_server *s, bfs_server_stats *stats);
/* Stops admitting, finishes resident queries and joins the server threads. */
void bfs_server_destroy(bfs_server *s);

/* Page-locked host memory for result buffers: copies into it run at full
 * link speed and overlap the next levels.  Any host memory works. */
bfs_status bfs_host_alloc_pinned(int64_t bytes, void **ptr);
void bfs_host_free_pinned(void *ptr);

#ifdef __cplusplus
}
#endif

#endif /* B40C_BFS_BATCH_H */
