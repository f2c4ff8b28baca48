/*
 * bfs_oracle.c -- CPU restatement of b40c's serial reference BFS and of the
 * test driver's scoring and validation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the reference the HIP traversal
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Every function cites the reference file:line it follows.  Parity pinning
 * (each is checked by tests/test_oracle.py against something other than a
 * retyped copy of itself):
 *   oracle_bfs            closed forms on grid lattices (d = i+j, d = i+j+k),
 *                         brute-force relaxation / Floyd-Warshall on tiny
 *                         graphs, hand-computed golden vectors, BFS invariants.
 *   oracle_edges_visited  closed form on grids (E_vis = E when connected),
 *                         hand-computed totals on golden graphs.
 *   oracle_compare        identity / single planted mismatch.
 *   oracle_check_preds    accepts the parent trees of several BFS orders,
 *                         rejects planted corruptions of each kind.
 *   oracle_level_counts   grid closed forms (number of lattice points with a
 *                         given coordinate sum).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec * 1e3 + (double) ts.tv_nsec * 1e-6;
}

/*
 * Serial FIFO breadth-first search.
 * Follows test/bfs/test_bfs.cu:514-554 (SimpleReferenceBfs): labels start at
 * -1, labels[src] = 0, a vertex is appended to the FIFO exactly when its label
 * is still -1, and it receives its parent's label + 1.  The reference's
 * std::deque is replaced by a flat array: every vertex is enqueued at most
 * once, so V slots always suffice and the visiting order is the same.
 *
 * Returns the search depth as the reference prints it (max label + 1,
 * test_bfs.cu:545-557), or -1 on invalid arguments.  *elapsed_ms receives the
 * wall-clock time of the traversal loop only (the scope of test_bfs.cu:529-555;
 * the reference used getrusage user time, b40c_test_util.h:503-522).
 * If preds is non-NULL it receives the FIFO parent of every vertex (-2 for the
 * source, -1 unvisited), the predecessor form of the labels
 * (b40c/graph/bfs/csr_problem.cuh:255-273 with MARK_PREDECESSORS).
 */
int32_t oracle_bfs(int32_t nodes, const int32_t *row_offsets, const int32_t *column_indices,
                   int32_t src, int32_t *labels, int32_t *preds, double *elapsed_ms)
{
    if (nodes <= 0 || src < 0 || src >= nodes || !row_offsets || !column_indices || !labels)
        return -1;
    int32_t *queue = (int32_t *) malloc(sizeof(int32_t) * (size_t) nodes);
    if (!queue) return -1;

    for (int32_t i = 0; i < nodes; i++) labels[i] = -1;
    if (preds)
        for (int32_t i = 0; i < nodes; i++) preds[i] = -1;
    labels[src] = 0;
    if (preds) preds[src] = -2;
    int32_t search_depth = 0;

    int64_t head = 0, tail = 0;
    queue[tail++] = src;

    double t0 = now_ms();
    while (head < tail) {
        int32_t dequeued = queue[head++];
        int32_t neighbor_dist = labels[dequeued] + 1;
        int32_t e_begin = row_offsets[dequeued];
        int32_t e_end = row_offsets[dequeued + 1];
        for (int32_t e = e_begin; e < e_end; e++) {
            int32_t neighbor = column_indices[e];
            if (labels[neighbor] == -1) {
                labels[neighbor] = neighbor_dist;
                if (preds) preds[neighbor] = dequeued;
                if (search_depth < neighbor_dist) search_depth = neighbor_dist;
                queue[tail++] = neighbor;
            }
        }
    }
    double t1 = now_ms();
    if (elapsed_ms) *elapsed_ms = t1 - t0;
    free(queue);
    return search_depth + 1;
}

/*
 * Traversed-edge accounting.  Follows test/bfs/test_bfs.cu:360-367: a vertex
 * with label > -1 is visited and contributes its full out-degree.
 * Returns edges_visited; *nodes_visited receives the visited-vertex count.
 */
int64_t oracle_edges_visited(int32_t nodes, const int32_t *row_offsets, const int32_t *labels,
                             int64_t *nodes_visited)
{
    int64_t ev = 0, nv = 0;
    for (int32_t i = 0; i < nodes; i++) {
        if (labels[i] > -1) {
            nv++;
            ev += (int64_t) row_offsets[i + 1] - row_offsets[i];
        }
    }
    if (nodes_visited) *nodes_visited = nv;
    return ev;
}

/*
 * Exact element-wise comparison.  Follows test/b40c_test_util.h:303-335
 * (CompareResults).  Returns -1 when equal, else the first mismatching index.
 */
int64_t oracle_compare(const int32_t *computed, const int32_t *reference, int64_t len)
{
    for (int64_t i = 0; i < len; i++)
        if (computed[i] != reference[i]) return i;
    return -1;
}

/*
 * Predecessor plausibility check.  Follows test/bfs/test_bfs.cu:387-433 and
 * tightens it as the predecessor contract requires:
 *   - a vertex with no parent (pred < 0) must not have a positive depth
 *     (test_bfs.cu:392-401); additionally pred == -2 exactly for depth 0 and
 *     pred == -1 exactly for unreached vertices;
 *   - the parent's depth is exactly one less (test_bfs.cu:403-410);
 *   - the parent lists the vertex in its adjacency row (test_bfs.cu:412-428),
 *     i.e. the parent is an in-neighbour.
 * depth[] are reference depths (from oracle_bfs).
 * Returns -1 if every vertex passes, else the first failing vertex.
 */
int64_t oracle_check_preds(int32_t nodes, const int32_t *row_offsets, const int32_t *column_indices,
                           const int32_t *depth, const int32_t *preds)
{
    for (int32_t v = 0; v < nodes; v++) {
        int32_t p = preds[v];
        int32_t d = depth[v];
        if (p < 0) {
            if (d > 0) return v;
            if (p == -2 && d != 0) return v;
            if (p == -1 && d != -1) return v;
            if (p != -1 && p != -2) return v;
            continue;
        }
        if (p >= nodes) return v;
        if (depth[p] + 1 != d) return v;
        int found = 0;
        for (int32_t e = row_offsets[p]; e < row_offsets[p + 1]; e++) {
            if (column_indices[e] == v) { found = 1; break; }
        }
        if (!found) return v;
    }
    return -1;
}

/*
 * Vertices discovered per depth.  Follows the "discovered" column of the
 * reference's work histogram (test/bfs/test_bfs.cu:264-290).  counts[] has
 * max_depth entries; returns the number of levels (max label + 1), or -1 if a
 * label exceeds max_depth - 1.
 */
int32_t oracle_level_counts(int32_t nodes, const int32_t *labels, int64_t *counts, int32_t max_depth)
{
    memset(counts, 0, sizeof(int64_t) * (size_t) max_depth);
    int32_t levels = 0;
    for (int32_t i = 0; i < nodes; i++) {
        int32_t d = labels[i];
        if (d < 0) continue;
        if (d >= max_depth) return -1;
        counts[d]++;
        if (d + 1 > levels) levels = d + 1;
    }
    return levels;
}
