// graph_builder.cpp -- host CSR construction and synthetic generators.
//
// Replaces the reference's host graph container and builders used by the BFS
// driver (b40c/graph/csr_graph.cuh:50-245, b40c/graph/builder/*.cuh).  The
// interface is declared in include/b40c_graph.h.  Everything here is host
// C++ with std::thread parallelism; results do not depend on the thread count
// because every random draw is a pure function of (seed, item index).
#include "b40c_graph.h"
#include "graph_builder_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// SplitMix64 finaliser (Steele, Lea, Flood: "Fast splittable pseudorandom
// number generators", OOPSLA 2014).
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based stream: the i-th stream of a seed is independent of how the
// work is split over threads.
struct Stream {
    uint64_t s;
    Stream(uint64_t seed, uint64_t index) : s(mix64(seed ^ mix64(index * kGolden + 0x632BE59BD9B4E019ull))) {}
    uint64_t next() { s += kGolden; return mix64(s); }
    // Unbiased-enough bounded draw (Lemire multiply-shift); bias < 2^-32 for
    // the ranges used here.
    uint64_t below(uint64_t range) { return (uint64_t) (((__uint128_t) next() * range) >> 64); }
};

int host_threads()
{
    const char *env = std::getenv("B40C_HOST_THREADS");
    int t = env ? std::atoi(env) : (int) std::thread::hardware_concurrency();
    if (t < 1) t = 1;
    if (t > 16) t = 16;
    return t;
}

template <typename F>
void parallel_for(int64_t n, F &&fn)
{
    int t = host_threads();
    if (n < 1 << 16 || t == 1) {
        fn(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    int64_t chunk = (n + t - 1) / t;
    for (int i = 0; i < t; i++) {
        int64_t b = i * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&fn, b, e, i] { fn(b, e, i); });
    }
    for (auto &th : pool) th.join();
}

int alloc_csr(int32_t nodes, int64_t edges, b40c_csr *out)
{
    out->nodes = nodes;
    out->edges = edges;
    out->row_offsets = (int32_t *) std::malloc(sizeof(int32_t) * ((size_t) nodes + 1));
    out->column_indices = (int32_t *) std::malloc(sizeof(int32_t) * (size_t) std::max<int64_t>(edges, 1));
    if (!out->row_offsets || !out->column_indices) {
        std::free(out->row_offsets);
        std::free(out->column_indices);
        out->row_offsets = out->column_indices = nullptr;
        return 2;
    }
    return 0;
}

// Builds a CSR from (src, dst) pairs.  Every row is sorted by column index;
// with dedup, repeated entries collapse to one; with drop_loops, (v, v) is
// removed; with symmetrize, (v, u) is added for every (u, v).
int csr_from_pairs(int32_t nodes, int64_t num_pairs, const int32_t *src, const int32_t *dst,
                   bool symmetrize, bool dedup, bool drop_loops, b40c_csr *out)
{
    if (nodes <= 0 || num_pairs < 0) return 1;
    std::vector<std::atomic<int32_t>> deg(nodes);
    for (auto &d : deg) d.store(0, std::memory_order_relaxed);

    parallel_for(num_pairs, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; i++) {
            int32_t u = src[i], v = dst[i];
            if (drop_loops && u == v) continue;
            deg[u].fetch_add(1, std::memory_order_relaxed);
            if (symmetrize) deg[v].fetch_add(1, std::memory_order_relaxed);
        }
    });

    std::vector<int64_t> start((size_t) nodes + 1);
    start[0] = 0;
    for (int32_t v = 0; v < nodes; v++) start[v + 1] = start[v] + deg[v].load(std::memory_order_relaxed);
    int64_t total = start[nodes];

    int32_t *tmp = (int32_t *) std::malloc(sizeof(int32_t) * (size_t) std::max<int64_t>(total, 1));
    if (!tmp) return 2;
    std::vector<std::atomic<int64_t>> cursor(nodes);
    for (int32_t v = 0; v < nodes; v++) cursor[v].store(start[v], std::memory_order_relaxed);

    parallel_for(num_pairs, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; i++) {
            int32_t u = src[i], v = dst[i];
            if (drop_loops && u == v) continue;
            tmp[cursor[u].fetch_add(1, std::memory_order_relaxed)] = v;
            if (symmetrize) tmp[cursor[v].fetch_add(1, std::memory_order_relaxed)] = u;
        }
    });

    // Sort (and optionally deduplicate) each row; record the new row lengths.
    std::vector<int32_t> newdeg(nodes);
    parallel_for(nodes, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; v++) {
            int32_t *row = tmp + start[v];
            int64_t len = start[v + 1] - start[v];
            std::sort(row, row + len);
            if (dedup) len = std::unique(row, row + len) - row;
            newdeg[v] = (int32_t) len;
        }
    });

    int64_t edges = 0;
    for (int32_t v = 0; v < nodes; v++) edges += newdeg[v];
    if (edges > INT32_MAX) {
        std::free(tmp);
        return 3;
    }
    int rc = alloc_csr(nodes, edges, out);
    if (rc) {
        std::free(tmp);
        return rc;
    }
    int64_t acc = 0;
    for (int32_t v = 0; v < nodes; v++) {
        out->row_offsets[v] = (int32_t) acc;
        acc += newdeg[v];
    }
    out->row_offsets[nodes] = (int32_t) acc;
    parallel_for(nodes, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; v++)
            std::memcpy(out->column_indices + out->row_offsets[v], tmp + start[v], sizeof(int32_t) * newdeg[v]);
    });
    std::free(tmp);
    return 0;
}

}  // namespace

// Seeded vertex scramble of the R-MAT builders (host and device): a
// Fisher-Yates shuffle of 0 .. 2^scale - 1 driven by one counter stream.
void b40c_rmat_permutation(int32_t scale, uint64_t seed, std::vector<int32_t> *perm)
{
    const int32_t nodes = (int32_t) (1u << scale);
    perm->resize(nodes);
    for (int32_t v = 0; v < nodes; v++) (*perm)[v] = v;
    Stream st(seed ^ 0x5CA3B1E5CA3B1E5Cull, 0);
    for (int32_t v = nodes - 1; v > 0; v--) {
        int32_t j = (int32_t) st.below((uint64_t) v + 1);
        std::swap((*perm)[v], (*perm)[j]);
    }
}

extern "C" {

// Neighbour order (k-1, k+1, j-1, j+1) follows b40c/graph/builder/grid2d.cuh:78-100;
// the reference's trailing self-loop (grid2d.cuh:102-104) is not emitted.
int b40c_graph_grid2d(int32_t width, b40c_csr *out)
{
    if (!out || width <= 0 || (int64_t) width * width > INT32_MAX) return 1;
    int64_t w = width;
    int32_t nodes = (int32_t) (w * w);
    int64_t edges = w >= 2 ? 4 * w * (w - 1) : 0;  // 2 * (horizontal + vertical links)
    if (edges > INT32_MAX) return 3;
    int rc = alloc_csr(nodes, edges, out);
    if (rc) return rc;
    int64_t total = 0;
    for (int64_t j = 0; j < w; j++) {
        for (int64_t k = 0; k < w; k++) {
            int64_t me = j * w + k;
            out->row_offsets[me] = (int32_t) total;
            if (k - 1 >= 0) out->column_indices[total++] = (int32_t) (me - 1);
            if (k + 1 < w) out->column_indices[total++] = (int32_t) (me + 1);
            if (j - 1 >= 0) out->column_indices[total++] = (int32_t) (me - w);
            if (j + 1 < w) out->column_indices[total++] = (int32_t) (me + w);
        }
    }
    out->row_offsets[nodes] = (int32_t) total;
    return 0;
}

// Neighbour order (k-1, k+1, j-1, j+1, i-1, i+1) follows
// b40c/graph/builder/grid3d.cuh:73-114; no self-loop (grid3d.cuh:116-118).
int b40c_graph_grid3d(int32_t width, b40c_csr *out)
{
    if (!out || width <= 0 || (int64_t) width * width * width > INT32_MAX) return 1;
    int64_t w = width;
    int32_t nodes = (int32_t) (w * w * w);
    int64_t edges = w >= 2 ? 6 * w * w * (w - 1) : 0;  // 2 * 3 * w^2 * (w-1) links
    if (edges > INT32_MAX) return 3;
    int rc = alloc_csr(nodes, edges, out);
    if (rc) return rc;
    int64_t plane = w * w;
    // Row lengths have a closed form, so planes can be filled in parallel:
    // a vertex has deg(k) + deg(j) + deg(i) neighbours with deg(a) = [a>0] + [a<w-1],
    // so plane ii holds 2*w*S + w*w*deg(ii) entries with S = sum_a deg(a) = 2(w-1).
    auto plane_start = [&](int64_t i) -> int64_t {
        auto deg = [&](int64_t a) -> int64_t { return (a > 0) + (a < w - 1); };
        int64_t S = 2 * (w - 1), s = 0;
        for (int64_t ii = 0; ii < i; ii++) s += 2 * w * S + plane * deg(ii);
        return s;
    };
    parallel_for(w, [&](int64_t ib, int64_t ie, int) {
        for (int64_t i = ib; i < ie; i++) {
            int64_t total = plane_start(i);
            for (int64_t j = 0; j < w; j++) {
                for (int64_t k = 0; k < w; k++) {
                    int64_t me = i * plane + j * w + k;
                    out->row_offsets[me] = (int32_t) total;
                    if (k - 1 >= 0) out->column_indices[total++] = (int32_t) (me - 1);
                    if (k + 1 < w) out->column_indices[total++] = (int32_t) (me + 1);
                    if (j - 1 >= 0) out->column_indices[total++] = (int32_t) (me - w);
                    if (j + 1 < w) out->column_indices[total++] = (int32_t) (me + w);
                    if (i - 1 >= 0) out->column_indices[total++] = (int32_t) (me - plane);
                    if (i + 1 < w) out->column_indices[total++] = (int32_t) (me + plane);
                }
            }
        }
    });
    out->row_offsets[nodes] = (int32_t) edges;
    return 0;
}

// Graph500-style R-MAT: each of edge_factor * 2^scale samples descends `scale`
// levels of the adjacency quadtree, picking quadrant (0,0), (0,1), (1,0),
// (1,1) with probabilities a, b, c, 1-a-b-c.  Unlike the reference's
// GTgraph-style builder (b40c/graph/builder/rmat.cuh:64-116, 167-168) there is
// no per-level parameter noise and ids are 0-based.  With `scramble`, vertex
// ids pass through a seeded random permutation (as Graph500 prescribes) so the
// hubs are not clustered at low ids.  The result is symmetrised with
// self-loops and duplicate edges removed.
int b40c_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                    uint64_t seed, int scramble, b40c_csr *out)
{
    if (!out || scale < 1 || scale > 30 || edge_factor < 1) return 1;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return 1;
    int32_t nodes = (int32_t) (1u << scale);
    int64_t samples = (int64_t) edge_factor << scale;
    if (2 * samples > (int64_t) INT32_MAX * 2) return 3;

    const uint64_t two32 = 1ull << 32;
    const uint64_t t_a = (uint64_t) (a * two32);
    const uint64_t t_ab = (uint64_t) ((a + b) * two32);
    const uint64_t t_abc = (uint64_t) ((a + b + c) * two32);

    std::vector<int32_t> perm;
    if (scramble) b40c_rmat_permutation(scale, seed, &perm);

    int32_t *src = (int32_t *) std::malloc(sizeof(int32_t) * (size_t) samples);
    int32_t *dst = (int32_t *) std::malloc(sizeof(int32_t) * (size_t) samples);
    if (!src || !dst) {
        std::free(src);
        std::free(dst);
        return 2;
    }
    parallel_for(samples, [&](int64_t bgn, int64_t end, int) {
        for (int64_t e = bgn; e < end; e++) {
            Stream st(seed, (uint64_t) e);
            uint32_t u = 0, v = 0;
            uint64_t bits = 0;
            for (int32_t lvl = 0; lvl < scale; lvl++) {
                if ((lvl & 1) == 0) bits = st.next();
                uint64_t r = (lvl & 1) ? (bits >> 32) : (bits & 0xFFFFFFFFull);
                uint32_t ub, vb;
                if (r < t_a) { ub = 0; vb = 0; }
                else if (r < t_ab) { ub = 0; vb = 1; }
                else if (r < t_abc) { ub = 1; vb = 0; }
                else { ub = 1; vb = 1; }
                u = (u << 1) | ub;
                v = (v << 1) | vb;
            }
            if (scramble) {
                u = (uint32_t) perm[u];
                v = (uint32_t) perm[v];
            }
            src[e] = (int32_t) u;
            dst[e] = (int32_t) v;
        }
    });
    int rc = csr_from_pairs(nodes, samples, src, dst, true, true, true, out);
    std::free(src);
    std::free(dst);
    return rc;
}

// Uniform G(n, m): exactly m distinct undirected edges {u, v}, u != v, drawn
// uniformly; stored symmetrised (2m directed entries).
int b40c_graph_erdos_renyi(int32_t n, int64_t m, uint64_t seed, b40c_csr *out)
{
    if (!out || n < 2 || m < 0) return 1;
    if ((double) m > (double) n * (n - 1) / 2.0) return 1;
    if (2 * m > INT32_MAX) return 3;
    std::vector<uint64_t> keys;
    uint64_t counter = 0;
    while ((int64_t) keys.size() < m) {
        int64_t need = m - (int64_t) keys.size();
        std::vector<uint64_t> batch((size_t) need);
        uint64_t base = counter;
        parallel_for(need, [&](int64_t b, int64_t e, int) {
            for (int64_t i = b; i < e; i++) {
                Stream st(seed, base + (uint64_t) i);
                uint32_t u = (uint32_t) st.below((uint64_t) n);
                uint32_t v = (uint32_t) st.below((uint64_t) n);
                if (u == v) { batch[i] = ~0ull; continue; }
                if (u > v) std::swap(u, v);
                batch[i] = ((uint64_t) u << 32) | v;
            }
        });
        counter += (uint64_t) need;
        std::sort(batch.begin(), batch.end());
        batch.erase(std::unique(batch.begin(), batch.end()), batch.end());
        if (!batch.empty() && batch.back() == ~0ull) batch.pop_back();
        std::vector<uint64_t> merged;
        merged.reserve(keys.size() + batch.size());
        std::merge(keys.begin(), keys.end(), batch.begin(), batch.end(), std::back_inserter(merged));
        merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
        keys.swap(merged);
    }
    std::vector<int32_t> src((size_t) m), dst((size_t) m);
    for (int64_t i = 0; i < m; i++) {
        src[i] = (int32_t) (keys[i] >> 32);
        dst[i] = (int32_t) (keys[i] & 0xFFFFFFFFu);
    }
    return csr_from_pairs(n, m, src.data(), dst.data(), true, true, true, out);
}

// Random multigraph, as the reference's BuildRandomGraph
// (b40c/graph/builder/random.cuh:53-98): m directed (u, v) pairs drawn
// uniformly with loops and repeats kept; with `undirected` every pair also
// contributes (v, u).
int b40c_graph_random(int32_t n, int64_t m, uint64_t seed, int undirected, b40c_csr *out)
{
    if (!out || n < 1 || m < 0) return 1;
    if ((undirected ? 2 : 1) * m > INT32_MAX) return 3;
    std::vector<int32_t> src((size_t) m), dst((size_t) m);
    parallel_for(m, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; i++) {
            Stream st(seed, (uint64_t) i);
            src[i] = (int32_t) st.below((uint64_t) n);
            dst[i] = (int32_t) st.below((uint64_t) n);
        }
    });
    return csr_from_pairs(n, m, src.data(), dst.data(), undirected != 0, false, false, out);
}

int b40c_graph_from_coo(int32_t nodes, int64_t num_pairs, const int32_t *src, const int32_t *dst,
                        int symmetrize, int dedup, int drop_self_loops, b40c_csr *out)
{
    if (!out || nodes <= 0 || num_pairs < 0 || (num_pairs > 0 && (!src || !dst))) return 1;
    for (int64_t i = 0; i < num_pairs; i++)
        if (src[i] < 0 || src[i] >= nodes || dst[i] < 0 || dst[i] >= nodes) return 1;
    return csr_from_pairs(nodes, num_pairs, src, dst, symmetrize != 0, dedup != 0, drop_self_loops != 0, out);
}

int b40c_graph_sample_sources(const b40c_csr *g, int32_t count, uint64_t seed, int32_t *out)
{
    if (!g || !out || count < 0 || !g->row_offsets) return 1;
    std::vector<int32_t> cand;
    cand.reserve(g->nodes);
    for (int32_t v = 0; v < g->nodes; v++)
        if (g->row_offsets[v + 1] > g->row_offsets[v]) cand.push_back(v);
    if ((int64_t) cand.size() < count) return 1;
    Stream st(seed ^ 0x50C4CE5ull, 0);
    int64_t k = (int64_t) cand.size();
    for (int32_t i = 0; i < count; i++) {
        int64_t j = i + (int64_t) st.below((uint64_t) (k - i));
        std::swap(cand[i], cand[j]);
        out[i] = cand[i];
    }
    return 0;
}

void b40c_graph_free(b40c_csr *g)
{
    if (!g) return;
    std::free(g->row_offsets);
    std::free(g->column_indices);
    g->row_offsets = g->column_indices = nullptr;
    g->nodes = 0;
    g->edges = 0;
}

}  // extern "C"
