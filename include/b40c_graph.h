/*
 * b40c_graph.h -- host CSR container and synthetic graph builders (C ABI).
 *
 * Exported by libb40c_hip.so.  These replace the reference's host-side graph
 * container and generators used by its BFS driver:
 *   b40c_csr                 <- CsrGraph<VertexId,Value,SizeT>
 *                               (b40c/graph/csr_graph.cuh:50-245), int32 only
 *   b40c_graph_grid2d        <- builder::BuildGrid2dGraph (b40c/graph/builder/grid2d.cuh:53-114)
 *                               same neighbour order, but without the self-loop
 *   b40c_graph_grid3d        <- builder::BuildGrid3dGraph (b40c/graph/builder/grid3d.cuh:53-129)
 *                               same neighbour order, but without the self-loop
 *   b40c_graph_rmat          <- builder::BuildRmatGraph (b40c/graph/builder/rmat.cuh:127-206)
 *                               redesigned: Graph500-style (no parameter noise,
 *                               0-based, seeded vertex scramble), symmetrised,
 *                               self-loops and duplicates removed
 *   b40c_graph_erdos_renyi   <- (new) uniform G(n, m): exactly m distinct undirected edges
 *   b40c_graph_random        <- builder::BuildRandomGraph (b40c/graph/builder/random.cuh:53-98)
 *                               keeps loops and multi-edges like the reference
 *   b40c_graph_from_coo      <- CsrGraph::FromCoo (b40c/graph/csr_graph.cuh:108-152)
 *   b40c_graph_sample_sources<- builder::RandomNode (b40c/graph/builder/utils.cuh:55-61),
 *                               restricted to vertices of nonzero degree
 *
 * All builders are deterministic for a given seed, independent of the number
 * of host threads used.  Rows of builders that deduplicate are sorted by
 * column index.  Return value: 0 on success, nonzero on failure (invalid
 * argument = 1, out of memory = 2, size overflow of int32 offsets = 3).
 */
#ifndef B40C_GRAPH_H
#define B40C_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct b40c_csr {
    int32_t nodes;            /* V */
    int64_t edges;            /* E: number of directed CSR entries */
    int32_t *row_offsets;     /* [V + 1], malloc-owned by the library */
    int32_t *column_indices;  /* [E],     malloc-owned by the library */
} b40c_csr;

int b40c_graph_grid2d(int32_t width, b40c_csr *out);
int b40c_graph_grid3d(int32_t width, b40c_csr *out);
int b40c_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                    uint64_t seed, int scramble, b40c_csr *out);
int b40c_graph_erdos_renyi(int32_t n, int64_t m, uint64_t seed, b40c_csr *out);
int b40c_graph_random(int32_t n, int64_t m, uint64_t seed, int undirected, b40c_csr *out);
int b40c_graph_from_coo(int32_t nodes, int64_t num_pairs, const int32_t *src, const int32_t *dst,
                        int symmetrize, int dedup, int drop_self_loops, b40c_csr *out);
/* Draws `count` distinct sources uniformly from {v : deg(v) > 0}. */
int b40c_graph_sample_sources(const b40c_csr *g, int32_t count, uint64_t seed, int32_t *out);
void b40c_graph_free(b40c_csr *g);

/*
 * Graph files, as directed (src, dst) pairs, 0-based, in file order; build
 * the CSR with b40c_graph_from_coo or b40c_device_csr_from_coo.  Same
 * grammar, id base, count checks and undirected doubling as the reference's
 * readers:
 *   b40c_read_dimacs  <- ReadDimacsStream (b40c/graph/builder/dimacs.cuh:53-225)
 *   b40c_read_metis   <- ReadMetisStream  (b40c/graph/builder/metis.cuh:51-176)
 *   b40c_read_market  <- ReadMarketStream (b40c/graph/builder/market.cuh:49-164);
 *                        an entry "i j" is the arc j -> i, as there
 * Return codes as above, plus 5 = the file cannot be read; a malformed
 * file is 1.
 */
typedef struct b40c_coo {
    int32_t nodes;
    int64_t pairs;
    int32_t *src;   /* [pairs], malloc-owned by the library */
    int32_t *dst;   /* [pairs] */
} b40c_coo;

int b40c_read_dimacs(const char *path, int undirected, b40c_coo *out);
int b40c_read_metis(const char *path, b40c_coo *out);
int b40c_read_market(const char *path, int undirected, b40c_coo *out);
void b40c_coo_free(b40c_coo *c);

/*
 * addition: CSR construction on the device, the step before the search.
 * The reference builds every CSR on the host (CsrGraph::FromCoo,
 * b40c/graph/csr_graph.cuh:108-152, after the builders of
 * b40c/graph/builder/).  These sort the tuples on the GPU (LSD radix
 * sort of (row, column) keys, single-pass chained scans, stable compaction
 * for dedup) and leave the CSR in device memory, rows sorted by column: the
 * same arrays as the host builders above for the same input.
 *   b40c_device_csr_from_coo <- b40c_graph_from_coo on host pairs
 *   b40c_device_csr_rmat     <- b40c_graph_rmat (same samples, drawn on the device)
 *   b40c_device_csr_transpose: the in-neighbour CSR of a device CSR
 * Return codes as above, plus 4 = device error (no device, HIP failure).
 */
typedef struct b40c_device_csr {
    int32_t device;
    int32_t nodes;
    int64_t edges;
    int32_t *row_offsets;     /* device [V + 1], owned (b40c_device_csr_free) */
    int32_t *column_indices;  /* device [E] */
    int32_t symmetric;        /* 1 when built symmetrised: every (u, v) has (v, u) */
    double build_ms;          /* device time of the construction (hipEvents) */
} b40c_device_csr;

int b40c_device_csr_from_coo(int32_t device, int32_t nodes, int64_t num_pairs, const int32_t *h_src,
                             const int32_t *h_dst, int symmetrize, int dedup, int drop_self_loops,
                             b40c_device_csr *out);
int b40c_device_csr_rmat(int32_t device, int32_t scale, int32_t edge_factor, double a, double b, double c,
                         uint64_t seed, int scramble, b40c_device_csr *out);
int b40c_device_csr_transpose(const b40c_device_csr *g, b40c_device_csr *out);
int b40c_device_csr_to_host(const b40c_device_csr *g, b40c_csr *out);
void b40c_device_csr_free(b40c_device_csr *g);

#ifdef __cplusplus
}
#endif

#endif /* B40C_GRAPH_H */
