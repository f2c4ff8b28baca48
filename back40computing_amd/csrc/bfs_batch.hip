// bfs_batch.hip -- host side of batched BFS and of the continuous-batching
// server (C ABI in include/b40c_bfs_batch.h).
//
// Graph pool: device-resident CSR graphs (plus the in-neighbour CSR of
// asymmetric graphs for bottom-up levels) and a device table of them.
// Batch: an arena of per-query state (labels, visited mask, two vertex
// queues) carved per enact, a table of BatchSlot descriptors, and either one
// persistent launch or replays of a hipGraph holding `levels_per_graph`
// (level kernel, advance kernel) pairs.
// Server: the same slot table with a fixed number of slots; a host thread
// admits queued requests into free slots between graph replays while the
// other slots keep running, retires finished queries (labels copied to the
// caller's buffer) and frees their memory.
#include "b40c_bfs_batch.h"
#include "b40c_graph.h"
#include "bfs_batch.hpp"
#include "bfs_batch_internal.hpp"
#include "bfs_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace b40c_hip;

namespace {

#define HIP_OK(expr, what)                                  \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return host_hip_fail(e_, what); \
    } while (0)

struct PoolGraph {
    int *row = nullptr, *col = nullptr, *rrow = nullptr, *rcol = nullptr;
    int32_t nodes = 0;
    int64_t edges = 0;
};

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

constexpr unsigned kBatchHubCapacity = 1u << 18;  // grid-tier rows per level

// Per-query device state layout inside an arena (256-byte aligned pieces).
struct SlotLayout {
    size_t depth, visited, fbm0, fbm1, q0, q1, bytes;
};

SlotLayout slot_layout(int32_t nodes)
{
    SlotLayout l;
    size_t v = (size_t) std::max(nodes, 1);
    l.depth = 0;
    l.visited = align_up(4 * v, 256);
    size_t words = align_up(4 * ((v + 31) / 32), 256);
    l.fbm0 = l.visited + words;
    l.fbm1 = l.fbm0 + words;
    l.q0 = l.fbm1 + words;
    l.q1 = l.q0 + align_up(4 * v, 256);
    l.bytes = l.q1 + align_up(4 * v, 256);
    return l;
}

double wall_ms(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

double percentile(std::vector<double> v, double p)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double idx = p * (v.size() - 1);
    size_t lo = (size_t) idx, hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (idx - lo);
}

}  // namespace

struct bfs_graph_pool {
    int device = 0;
    std::vector<PoolGraph> graphs;
    BatchGraph *d_table = nullptr;
    size_t table_cap = 0;
    int64_t bytes = 0;
    int num_cus = 0;
    int wall_khz = 100000;
    std::mutex mu;
};

// Shared slot machinery of batches and servers.
struct SlotTable {
    int device = 0;
    int max_slots = 0;
    hipStream_t stream = nullptr;
    BatchSlot *d_slots = nullptr;
    BatchSlot *h_slots = nullptr;   // pinned mirror
    BatchControl *d_ctl = nullptr;
    int *d_list = nullptr;          // slot indices to seed
    int *d_alive = nullptr;
    int *h_alive = nullptr;         // pinned
    HubEntry *d_hubs = nullptr;     // two grid-tier lists of kBatchHubCapacity entries
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_levels = 0;
    int graph_grid = 0;
    BatchArgs graph_args{};         // arguments the captured graph was built with
    bfs_config cfg{};

    bfs_status init(int dev, int slots)
    {
        device = dev;
        max_slots = slots;
        HIP_OK(hipSetDevice(dev), "hipSetDevice");
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "create stream");
        HIP_OK(hipMalloc(&d_slots, sizeof(BatchSlot) * slots), "allocate slots");
        HIP_OK(hipHostMalloc(&h_slots, sizeof(BatchSlot) * slots, hipHostMallocDefault), "allocate pinned slots");
        HIP_OK(hipMalloc(&d_ctl, sizeof(BatchControl)), "allocate batch control");
        HIP_OK(hipMalloc(&d_list, sizeof(int) * slots), "allocate seed list");
        HIP_OK(hipMalloc(&d_alive, sizeof(int)), "allocate alive counter");
        HIP_OK(hipHostMalloc(&h_alive, sizeof(int), hipHostMallocDefault), "allocate pinned counter");
        HIP_OK(hipMalloc(&d_hubs, sizeof(HubEntry) * 2 * kBatchHubCapacity), "allocate grid-tier lists");
        memset(h_slots, 0, sizeof(BatchSlot) * slots);
        HIP_OK(hipMemcpy(d_slots, h_slots, sizeof(BatchSlot) * slots, hipMemcpyHostToDevice), "clear slots");
        HIP_OK(hipMemset(d_ctl, 0, sizeof(BatchControl)), "clear batch control");
        bfs_batch_config_default(&cfg);
        return BFS_OK;
    }

    void release()
    {
        (void) hipSetDevice(device);
        if (stream) (void) hipStreamSynchronize(stream);
        if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
        if (graph) (void) hipGraphDestroy(graph);
        if (d_slots) (void) hipFree(d_slots);
        if (h_slots) (void) hipHostFree(h_slots);
        if (d_ctl) (void) hipFree(d_ctl);
        if (d_list) (void) hipFree(d_list);
        if (d_alive) (void) hipFree(d_alive);
        if (h_alive) (void) hipHostFree(h_alive);
        if (d_hubs) (void) hipFree(d_hubs);
        if (stream) (void) hipStreamDestroy(stream);
        graph_exec = nullptr;
        graph = nullptr;
        stream = nullptr;
    }

    BatchArgs args(const BatchGraph *table, int nslots, int direction) const
    {
        BatchArgs a;
        std::memset(&a, 0, sizeof a);  // padding too: cached graphs compare argument bytes
        a.graphs = table;
        a.slots = d_slots;
        a.ctl = d_ctl;
        a.nslots = nslots;
        a.wave_t = cfg.wave_gather_threshold;
        a.cta_t = cfg.cta_gather_threshold;
        a.direction = direction;
        a.do_alpha = cfg.do_alpha;
        a.do_beta = cfg.do_beta;
        a.hubs[0] = d_hubs;
        a.hubs[1] = d_hubs + kBatchHubCapacity;
        a.hub_cap = kBatchHubCapacity;
        a.grid_t = cfg.grid_gather_threshold > 0 ? cfg.grid_gather_threshold : 0x7fffffff;
        a.grab_max = cfg.tile_grab_max > 0 ? cfg.tile_grab_max : 1;
        a.bu_tile = cfg.bu_tile_blocks > 0 ? cfg.bu_tile_blocks : 1;
        a.strided = cfg.grab_strided != 2;
        a.coherent_filter = cfg.claim_filter == 2;
        return a;
    }

    // Captures `levels` batch levels (level, hub and advance kernels) into a graph.
    bfs_status capture(const BatchArgs &a, int grid, int levels)
    {
        // the kernel arguments are baked into the graph: any change re-captures
        if (graph_exec && graph_levels == levels && graph_grid == grid &&
            std::memcmp(&graph_args, &a, sizeof a) == 0)
            return BFS_OK;
        if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
        if (graph) (void) hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        for (int l = 0; l < levels; l++) {
            hipError_t e = launch_batch_level(a, grid, l == levels - 1 ? d_alive : nullptr, stream);
            if (e != hipSuccess) {
                hipGraph_t dummy = nullptr;
                (void) hipStreamEndCapture(stream, &dummy);
                if (dummy) (void) hipGraphDestroy(dummy);
                return host_hip_fail(e, "capture level launches");
            }
        }
        HIP_OK(hipStreamEndCapture(stream, &graph), "end capture");
        HIP_OK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "instantiate level graph");
        graph_levels = levels;
        graph_grid = grid;
        graph_args = a;
        return BFS_OK;
    }
};

struct bfs_batch {
    bfs_graph_pool *pool = nullptr;
    SlotTable t;
    int max_queries = 0;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    int nq = 0;
    std::vector<double> latency_ms;
    std::vector<int32_t> nodes_of;
    BatchControl last{};   // control block of the last enact (level record)
};

namespace {

bfs_status refresh_table(bfs_graph_pool *p)
{
    std::vector<BatchGraph> host(p->graphs.size());
    for (size_t i = 0; i < p->graphs.size(); i++) {
        const PoolGraph &g = p->graphs[i];
        host[i] = BatchGraph{g.row, g.col, g.rrow, g.rcol, g.nodes, 0, (unsigned long long) g.edges};
    }
    if (host.size() > p->table_cap) {
        size_t cap = std::max<size_t>(64, 2 * host.size());
        BatchGraph *nt = nullptr;
        HIP_OK(hipMalloc(&nt, sizeof(BatchGraph) * cap), "allocate graph table");
        if (p->d_table) (void) hipFree(p->d_table);
        p->d_table = nt;
        p->table_cap = cap;
    }
    HIP_OK(hipMemcpy(p->d_table, host.data(), sizeof(BatchGraph) * host.size(), hipMemcpyHostToDevice),
           "upload graph table");
    return BFS_OK;
}

int batch_grid(const bfs_graph_pool *p, bool persistent, int blocks_per_cu)
{
    int occ = persistent ? batch_persistent_blocks_per_cu() : batch_level_blocks_per_cu();
    if (persistent) {
        // same co-residency margin as the single-source persistent kernel
        int cap = std::max(1, std::min(kMaxPersistentPerCu, occ / 2));
        return p->num_cus * std::min(cap, std::max(1, blocks_per_cu));
    }
    // level launches drain a shared tile counter, so the grid only sets how
    // many queries' working sets are in flight at once
    return p->num_cus * std::max(1, std::min(occ, std::max(1, blocks_per_cu)));
}

}  // namespace

extern "C" {

int64_t bfs_query_state_bytes(int32_t nodes) { return (int64_t) slot_layout(nodes).bytes; }

bfs_status bfs_batch_config_default(bfs_config *cfg)
{
    if (!cfg) return host_fail(BFS_ERR_INVALID_ARG, "null config");
    bfs_config_default(cfg);
    // Autotuned on the 256-query R-MAT pool (64 graphs, scale 14..20):
    // bottom-up tiles of up to 64 blocks (16384 consecutive vertices per
    // workgroup; levels with few blocks use smaller ones), two workgroups
    // per CU, large consecutive grabs, an early grid tier (rows of 256
    // neighbours and more) and a later turn to bottom-up than one search
    // alone takes (alpha 12).
    cfg->blocks_per_cu = 2;
    cfg->bu_tile_blocks = 64;
    cfg->tile_grab_max = 128;
    cfg->grab_strided = 1;
    cfg->grid_gather_threshold = 256;
    cfg->wave_gather_threshold = 64;
    cfg->cta_gather_threshold = 512;
    cfg->do_alpha = 12;
    cfg->do_beta = 48;
    return BFS_OK;
}

bfs_status bfs_pool_create(int32_t device, bfs_graph_pool **out)
{
    if (!out) return host_fail(BFS_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return host_fail(BFS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return host_fail(BFS_ERR_INVALID_ARG, "device out of range");
    HIP_OK(hipSetDevice(device), "hipSetDevice");
    bfs_graph_pool *p = new bfs_graph_pool();
    p->device = device;
    if (hipDeviceGetAttribute(&p->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        hipDeviceGetAttribute(&p->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) {
        delete p;
        return host_fail(BFS_ERR_HIP, "device attributes");
    }
    if (p->wall_khz <= 0) p->wall_khz = 100000;
    *out = p;
    return BFS_OK;
}

bfs_status bfs_pool_add_graph(bfs_graph_pool *p, const int32_t *h_row, const int32_t *h_col, int32_t nodes,
                              int64_t edges, int32_t *graph_id)
{
    if (!p || !graph_id) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    bfs_status vs = host_validate_csr(h_row, h_col, nodes, edges);
    if (vs != BFS_OK) return vs;
    std::lock_guard<std::mutex> lk(p->mu);
    HIP_OK(hipSetDevice(p->device), "hipSetDevice");
    PoolGraph g;
    g.nodes = nodes;
    g.edges = edges;
    size_t rb = sizeof(int) * ((size_t) nodes + 1), cb = sizeof(int) * (size_t) std::max<int64_t>(edges, 1);
    HIP_OK(hipMalloc(&g.row, rb), "allocate row offsets");
    HIP_OK(hipMalloc(&g.col, cb), "allocate column indices");
    HIP_OK(hipMemcpy(g.row, h_row, rb, hipMemcpyHostToDevice), "copy row offsets");
    if (edges > 0) HIP_OK(hipMemcpy(g.col, h_col, sizeof(int) * (size_t) edges, hipMemcpyHostToDevice), "copy columns");
    p->bytes += (int64_t) (rb + cb);
    if (host_csr_is_symmetric(h_row, h_col, nodes)) {
        g.rrow = g.row;
        g.rcol = g.col;
    } else {
        std::vector<int32_t> src((size_t) edges);
        for (int32_t v = 0; v < nodes; v++)
            for (int32_t k = h_row[v]; k < h_row[v + 1]; k++) src[k] = v;
        b40c_csr tr{};
        if (b40c_graph_from_coo(nodes, edges, h_col, src.data(), 0, 0, 0, &tr) != 0)
            return host_fail(BFS_ERR_OOM, "transpose construction failed");
        bool ok = hipMalloc(&g.rrow, rb) == hipSuccess && hipMalloc(&g.rcol, cb) == hipSuccess &&
                  hipMemcpy(g.rrow, tr.row_offsets, rb, hipMemcpyHostToDevice) == hipSuccess &&
                  (edges == 0 || hipMemcpy(g.rcol, tr.column_indices, sizeof(int) * (size_t) edges,
                                           hipMemcpyHostToDevice) == hipSuccess);
        b40c_graph_free(&tr);
        if (!ok) return host_fail(BFS_ERR_OOM, "allocate or copy the transpose");
        p->bytes += (int64_t) (rb + cb);
    }
    p->graphs.push_back(g);
    bfs_status st = refresh_table(p);
    if (st != BFS_OK) return st;
    *graph_id = (int32_t) p->graphs.size() - 1;
    return BFS_OK;
}

bfs_status bfs_pool_add_device_graph(bfs_graph_pool *p, const b40c_device_csr *dg, int32_t *graph_id)
{
    if (!p || !dg || !graph_id || !dg->row_offsets || dg->nodes <= 0 || dg->edges < 0 || dg->edges > INT32_MAX ||
        (dg->edges > 0 && !dg->column_indices))
        return host_fail(BFS_ERR_INVALID_ARG, "invalid device graph");
    if (dg->device != p->device) return host_fail(BFS_ERR_INVALID_ARG, "device graph lives on another device");
    std::lock_guard<std::mutex> lk(p->mu);
    HIP_OK(hipSetDevice(p->device), "hipSetDevice");
    PoolGraph g;
    g.nodes = dg->nodes;
    g.edges = dg->edges;
    size_t rb = sizeof(int) * ((size_t) g.nodes + 1), cb = sizeof(int) * (size_t) std::max<int64_t>(g.edges, 1);
    HIP_OK(hipMalloc(&g.row, rb), "allocate row offsets");
    HIP_OK(hipMalloc(&g.col, cb), "allocate column indices");
    HIP_OK(hipMemcpy(g.row, dg->row_offsets, rb, hipMemcpyDeviceToDevice), "copy row offsets");
    if (g.edges > 0)
        HIP_OK(hipMemcpy(g.col, dg->column_indices, sizeof(int) * (size_t) g.edges, hipMemcpyDeviceToDevice),
               "copy columns");
    p->bytes += (int64_t) (rb + cb);
    if (dg->symmetric) {
        g.rrow = g.row;
        g.rcol = g.col;
    } else {
        // the in-neighbour CSR, sorted on the device as well
        b40c_device_csr t{};
        if (b40c_device_csr_transpose(dg, &t) != 0) return host_fail(BFS_ERR_OOM, "device transpose failed");
        g.rrow = t.row_offsets;
        g.rcol = t.column_indices;
        p->bytes += (int64_t) (rb + cb);
    }
    p->graphs.push_back(g);
    bfs_status st = refresh_table(p);
    if (st != BFS_OK) return st;
    *graph_id = (int32_t) p->graphs.size() - 1;
    return BFS_OK;
}

bfs_status bfs_pool_graph_info(const bfs_graph_pool *p, int32_t graph_id, int32_t *nodes, int64_t *edges)
{
    if (!p || graph_id < 0 || graph_id >= (int32_t) p->graphs.size())
        return host_fail(BFS_ERR_INVALID_ARG, "unknown graph id");
    if (nodes) *nodes = p->graphs[graph_id].nodes;
    if (edges) *edges = p->graphs[graph_id].edges;
    return BFS_OK;
}

bfs_status bfs_pool_device_bytes(const bfs_graph_pool *p, int64_t *bytes)
{
    if (!p || !bytes) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    *bytes = p->bytes;
    return BFS_OK;
}

void bfs_pool_destroy(bfs_graph_pool *p)
{
    if (!p) return;
    (void) hipSetDevice(p->device);
    (void) hipDeviceSynchronize();
    for (auto &g : p->graphs) {
        if (g.rrow && g.rrow != g.row) (void) hipFree(g.rrow);
        if (g.rcol && g.rcol != g.col) (void) hipFree(g.rcol);
        (void) hipFree(g.row);
        (void) hipFree(g.col);
    }
    if (p->d_table) (void) hipFree(p->d_table);
    delete p;
}

bfs_status bfs_batch_create(bfs_graph_pool *pool, int32_t max_queries, bfs_batch **out)
{
    if (!pool || !out || max_queries < 1 || max_queries > kMaxSlots)
        return host_fail(BFS_ERR_INVALID_ARG, "invalid batch arguments (max_queries must be 1..256)");
    *out = nullptr;
    bfs_batch *b = new bfs_batch();
    b->pool = pool;
    b->max_queries = max_queries;
    bfs_status st = b->t.init(pool->device, max_queries);
    if (st != BFS_OK) {
        b->t.release();
        delete b;
        return st;
    }
    *out = b;
    return BFS_OK;
}

bfs_status bfs_batch_set_config(bfs_batch *b, const bfs_config *cfg)
{
    if (!b || !cfg) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    bfs_config c = *cfg, d;
    bfs_batch_config_default(&d);
    if (c.wave_gather_threshold <= 0) c.wave_gather_threshold = d.wave_gather_threshold;
    if (c.cta_gather_threshold <= 0) c.cta_gather_threshold = d.cta_gather_threshold;
    if (c.do_alpha <= 0) c.do_alpha = d.do_alpha;
    if (c.do_beta <= 0) c.do_beta = d.do_beta;
    if (c.blocks_per_cu <= 0) c.blocks_per_cu = d.blocks_per_cu;
    if (c.grid_gather_threshold == 0) c.grid_gather_threshold = d.grid_gather_threshold;
    if (c.tile_grab_max <= 0) c.tile_grab_max = d.tile_grab_max;
    if (c.bu_tile_blocks <= 0) c.bu_tile_blocks = d.bu_tile_blocks;
    if (c.grab_strided <= 0) c.grab_strided = d.grab_strided;
    if (c.claim_filter <= 0) c.claim_filter = d.claim_filter;
    b->t.cfg = c;
    return BFS_OK;
}

bfs_status bfs_batch_enact(bfs_batch *b, const int32_t *graph_ids, const int32_t *sources, int32_t nq,
                           int32_t direction_optimizing, int32_t mode, bfs_batch_stats *stats)
{
    if (!b || !graph_ids || !sources || nq < 1 || nq > b->max_queries || (mode != 0 && mode != 1))
        return host_fail(BFS_ERR_INVALID_ARG, "invalid batch enact arguments");
    bfs_graph_pool *p = b->pool;
    SlotTable &t = b->t;
    HIP_OK(hipSetDevice(p->device), "hipSetDevice");
    std::vector<SlotLayout> lay(nq);
    size_t total = 0;
    int max_nodes = 1;
    b->nodes_of.assign(nq, 0);
    for (int i = 0; i < nq; i++) {
        int gid = graph_ids[i];
        if (gid < 0 || gid >= (int) p->graphs.size()) return host_fail(BFS_ERR_INVALID_ARG, "unknown graph id");
        const PoolGraph &g = p->graphs[gid];
        if (sources[i] < 0 || sources[i] >= g.nodes) return host_fail(BFS_ERR_INVALID_ARG, "source out of range");
        lay[i] = slot_layout(g.nodes);
        total += lay[i].bytes;
        max_nodes = std::max(max_nodes, g.nodes);
        b->nodes_of[i] = g.nodes;
    }
    if (total > b->arena_bytes) {
        if (b->arena) (void) hipFree(b->arena);
        b->arena = nullptr;
        b->arena_bytes = 0;
        HIP_OK(hipMalloc(&b->arena, total), "allocate query state");
        b->arena_bytes = total;
    }
    size_t off = 0;
    std::vector<int> list(nq);
    for (int i = 0; i < nq; i++) {
        BatchSlot &s = t.h_slots[i];
        memset(&s, 0, sizeof(s));
        char *base = b->arena + off;
        s.graph = graph_ids[i];
        s.src = sources[i];
        s.depth = reinterpret_cast<int *>(base + lay[i].depth);
        s.visited = reinterpret_cast<unsigned *>(base + lay[i].visited);
        s.fbm[0] = reinterpret_cast<unsigned *>(base + lay[i].fbm0);
        s.fbm[1] = reinterpret_cast<unsigned *>(base + lay[i].fbm1);
        s.q[0] = reinterpret_cast<int *>(base + lay[i].q0);
        s.q[1] = reinterpret_cast<int *>(base + lay[i].q1);
        s.cap = (unsigned) std::max(p->graphs[graph_ids[i]].nodes, 1);
        off += lay[i].bytes;
        list[i] = i;
    }
    HIP_OK(hipMemcpyAsync(t.d_slots, t.h_slots, sizeof(BatchSlot) * nq, hipMemcpyHostToDevice, t.stream),
           "upload slots");
    HIP_OK(hipMemcpyAsync(t.d_list, list.data(), sizeof(int) * nq, hipMemcpyHostToDevice, t.stream), "upload list");
    HIP_OK(hipMemsetAsync(t.d_ctl, 0, sizeof(BatchControl), t.stream), "clear batch control");
    BatchArgs a = t.args(p->d_table, nq, direction_optimizing ? 1 : 0);
    HIP_OK(launch_batch_seed(a, t.d_list, nq, max_nodes, t.stream), "seed queries");
    HIP_OK(hipStreamSynchronize(t.stream), "synchronize seeding");

    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0), "event");
    HIP_OK(hipEventCreate(&e1), "event");
    int launches = 0;
    bfs_status st = BFS_OK;
    unsigned long long clock0 = 0;
    do {
        if (hipEventRecord(e0, t.stream) != hipSuccess) { st = host_fail(BFS_ERR_HIP, "record"); break; }
        if (mode == 0) {
            int grid = batch_grid(p, true, t.cfg.blocks_per_cu);
            hipError_t e = launch_batch_persistent(a, grid, t.stream);
            if (e != hipSuccess) { st = host_hip_fail(e, "launch persistent batch"); break; }
            launches++;
        } else {
            const int levels = 8;
            int grid = batch_grid(p, false, t.cfg.blocks_per_cu);
            st = t.capture(a, grid, levels);
            if (st != BFS_OK) break;
            while (true) {
                hipError_t e = hipGraphLaunch(t.graph_exec, t.stream);
                if (e != hipSuccess) { st = host_hip_fail(e, "launch level graph"); break; }
                launches += kBatchLevelLaunches * levels;
                e = hipMemcpyAsync(t.h_alive, t.d_alive, sizeof(int), hipMemcpyDeviceToHost, t.stream);
                if (e == hipSuccess) e = hipStreamSynchronize(t.stream);
                if (e != hipSuccess) { st = host_hip_fail(e, "read alive count"); break; }
                if (*t.h_alive == 0) break;
            }
            if (st != BFS_OK) break;
        }
        if (hipEventRecord(e1, t.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            st = host_fail(BFS_ERR_HIP, "record");
            break;
        }
    } while (0);
    float ms = 0;
    if (st == BFS_OK) (void) hipEventElapsedTime(&ms, e0, e1);
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    if (st != BFS_OK) return st;
    HIP_OK(hipMemcpy(t.h_slots, t.d_slots, sizeof(BatchSlot) * nq, hipMemcpyDeviceToHost), "read slots");
    BatchControl hc;
    HIP_OK(hipMemcpy(&hc, t.d_ctl, sizeof(BatchControl), hipMemcpyDeviceToHost), "read batch control");
    b->nq = nq;
    b->last = hc;
    b->latency_ms.assign(nq, 0.0);
    clock0 = ~0ull;
    for (int i = 0; i < nq; i++) clock0 = std::min(clock0, t.h_slots[i].t_begin);
    bool overflow = false;
    for (int i = 0; i < nq; i++) {
        const BatchSlot &s = t.h_slots[i];
        b->latency_ms[i] = s.t_end > clock0 ? (double) (s.t_end - clock0) / p->wall_khz : 0.0;
        overflow |= s.overflow != 0;
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->elapsed_ms = ms;
        stats->queries = nq;
        stats->levels = hc.level;
        stats->launches = launches;
        stats->latency_p50_ms = percentile(b->latency_ms, 0.50);
        stats->latency_p99_ms = percentile(b->latency_ms, 0.99);
        stats->latency_max_ms = percentile(b->latency_ms, 1.0);
    }
    if (overflow) return host_fail(BFS_ERR_OVERFLOW, "query frontier overflow");
    return BFS_OK;
}

bfs_status bfs_batch_query_latency(const bfs_batch *b, double *ms)
{
    if (!b || !ms) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < b->nq; i++) ms[i] = b->latency_ms[i];
    return BFS_OK;
}

bfs_status bfs_batch_extract_depths(const bfs_batch *b, int32_t query, int32_t *h_depths)
{
    if (!b || !h_depths || query < 0 || query >= b->nq) return host_fail(BFS_ERR_INVALID_ARG, "invalid query");
    HIP_OK(hipSetDevice(b->pool->device), "hipSetDevice");
    HIP_OK(hipMemcpy(h_depths, b->t.h_slots[query].depth, sizeof(int) * (size_t) b->nodes_of[query],
                     hipMemcpyDeviceToHost),
           "copy depths");
    return BFS_OK;
}

bfs_status bfs_batch_level_log(const bfs_batch *b, int32_t max_levels, uint32_t *td_tiles, uint32_t *bu_tiles,
                               double *start_us, int32_t *levels)
{
    if (!b || !levels || max_levels < 0) return host_fail(BFS_ERR_INVALID_ARG, "invalid argument");
    const BatchControl &h = b->last;
    int k = 0;
    for (; k < kLevelLog && k < max_levels && k < h.level; k++) {
        if (td_tiles) td_tiles[k] = h.level_td_tiles[k];
        if (bu_tiles) bu_tiles[k] = h.level_bu_tiles[k];
        if (start_us) start_us[k] = 1000.0 * (double) (h.level_clock[k] - h.level_clock[0]) / b->pool->wall_khz;
    }
    *levels = k;
    return BFS_OK;
}

void bfs_batch_destroy(bfs_batch *b)
{
    if (!b) return;
    b->t.release();
    if (b->arena) (void) hipFree(b->arena);
    delete b;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Continuous-batching server
// ---------------------------------------------------------------------------
namespace {

struct Request {
    int64_t ticket = 0;
    int32_t graph = 0, src = 0;
    int32_t *out = nullptr;
    double submit_ms = 0, admit_ms = 0, finish_ms = 0;
    int32_t levels = 0;
    bfs_status status = BFS_OK;
    bool done = false;
    int slot = -1;
    size_t offset = 0, bytes = 0;
};

// First-fit allocator over [0, size) with coalescing frees.
struct Arena {
    std::map<size_t, size_t> free_list;  // offset -> bytes
    void reset(size_t size)
    {
        free_list.clear();
        if (size) free_list[0] = size;
    }
    bool alloc(size_t bytes, size_t *offset)
    {
        for (auto it = free_list.begin(); it != free_list.end(); ++it) {
            if (it->second >= bytes) {
                *offset = it->first;
                size_t rest = it->second - bytes, at = it->first + bytes;
                free_list.erase(it);
                if (rest) free_list[at] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t offset, size_t bytes)
    {
        auto it = free_list.emplace(offset, bytes).first;
        auto next = std::next(it);
        if (next != free_list.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_list.erase(next);
        }
        if (it != free_list.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_list.erase(it);
            }
        }
    }
};

}  // namespace

// A finished query whose labels are being copied out; its slot and memory
// are reused only after the copy completes.
struct DrainItem {
    int slot;
    int64_t ticket;
    const int *d_depth;
    size_t bytes;
    int32_t *out;
    int levels;
    bfs_status status;
};

struct bfs_server {
    bfs_graph_pool *pool = nullptr;
    bfs_server_config cfg{};
    SlotTable t;
    BatchGraph *d_table = nullptr;   // snapshot of the pool's graphs at creation
    std::vector<PoolGraph> graphs;
    char *d_arena = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Request> reqs;
    std::vector<int64_t> slot_ticket;
    std::vector<char> slot_draining;
    int resident = 0;                // admitted and not yet delivered
    int running = 0;                 // admitted and not finished on the device
    int draining = 0;                // finished, labels being copied out
    int64_t next_ticket = 1;
    bool stop = false, drain_stop = false;
    bfs_status fatal = BFS_OK;
    bfs_server_stats stats{};
    std::chrono::steady_clock::time_point t0;
    std::chrono::steady_clock::time_point busy_start;  // when resident last left 0
    std::thread thread, drainer;
    std::deque<DrainItem> drain_q;
    std::condition_variable cv_drain;
    hipStream_t copy_stream = nullptr;

    void run();
    void drain();
};

// Result-copy thread: copies each finished query's labels to its caller's
// buffer on a stream of its own, then delivers the result and frees the slot
// and its memory, while the server thread keeps replaying levels.
void bfs_server::drain()
{
    (void) hipSetDevice(pool->device);
    while (true) {
        DrainItem it;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_drain.wait(lk, [&] { return drain_stop || !drain_q.empty(); });
            if (drain_q.empty()) return;
            it = drain_q.front();
            drain_q.pop_front();
        }
        auto c0 = std::chrono::steady_clock::now();
        bfs_status rs = it.status;
        if (it.out && rs == BFS_OK) {
            hipError_t e = hipMemcpyAsync(it.out, it.d_depth, it.bytes, hipMemcpyDeviceToHost, copy_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(copy_stream);
            if (e != hipSuccess) rs = host_hip_fail(e, "copy result");
        }
        double cms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
        {
            std::lock_guard<std::mutex> lk(mu);
            Request &r = reqs[it.ticket];
            r.levels = it.levels;
            r.status = rs;
            r.finish_ms = wall_ms(t0);
            r.done = true;
            arena.release(r.offset, r.bytes);
            slot_ticket[it.slot] = -1;
            slot_draining[it.slot] = 0;
            draining--;
            resident--;
            stats.completed++;
            stats.copy_ms += cms;
            if (resident == 0) stats.busy_ms += wall_ms(busy_start);
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

void bfs_server::run()
{
    (void) hipSetDevice(pool->device);
    BatchArgs a = t.args(d_table, cfg.max_slots, cfg.direction_optimizing ? 1 : 0);
    int grid = batch_grid(pool, false, t.cfg.blocks_per_cu);
    bfs_status st = t.capture(a, grid, cfg.levels_per_tick);
    std::vector<int> admitted;
    while (true) {
        admitted.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            while (true) {
                if (st != BFS_OK) fatal = st;
                if (fatal != BFS_OK) {
                    // fail everything queued or running; draining results
                    // are delivered by the copy thread
                    for (auto &kv : reqs) {
                        Request &r = kv.second;
                        bool in_drain = r.slot >= 0 && slot_draining[r.slot] && slot_ticket[r.slot] == r.ticket;
                        if (!r.done && !in_drain) {
                            r.done = true;
                            r.status = fatal;
                        }
                    }
                    pending.clear();
                    cv_done.notify_all();
                    return;
                }
                // Admission: free slot and room in the HBM budget, in
                // arrival order.
                for (int s = 0; s < cfg.max_slots && !pending.empty(); s++) {
                    if (slot_ticket[s] >= 0) continue;
                    Request &r = reqs[pending.front()];
                    const PoolGraph &g = graphs[r.graph];
                    SlotLayout lay = slot_layout(g.nodes);
                    size_t off = 0;
                    if (!arena.alloc(lay.bytes, &off)) break;
                    r.offset = off;
                    r.bytes = lay.bytes;
                    r.slot = s;
                    r.admit_ms = wall_ms(t0);
                    slot_ticket[s] = r.ticket;
                    pending.pop_front();
                    BatchSlot &bs = t.h_slots[s];
                    memset(&bs, 0, sizeof(bs));
                    char *base = d_arena + off;
                    bs.graph = r.graph;
                    bs.src = r.src;
                    bs.depth = reinterpret_cast<int *>(base + lay.depth);
                    bs.visited = reinterpret_cast<unsigned *>(base + lay.visited);
                    bs.fbm[0] = reinterpret_cast<unsigned *>(base + lay.fbm0);
                    bs.fbm[1] = reinterpret_cast<unsigned *>(base + lay.fbm1);
                    bs.q[0] = reinterpret_cast<int *>(base + lay.q0);
                    bs.q[1] = reinterpret_cast<int *>(base + lay.q1);
                    bs.cap = (unsigned) std::max(g.nodes, 1);
                    admitted.push_back(s);
                    if (resident == 0) busy_start = std::chrono::steady_clock::now();
                    resident++;
                    running++;
                    stats.admitted++;
                    stats.max_resident = std::max<int64_t>(stats.max_resident, resident);
                }
                if (running > 0) break;
                if (stop && pending.empty() && draining == 0) return;
                // nothing to run: wait for a submission or a freed slot
                cv_work.wait(lk);
            }
        }
        auto tick0 = std::chrono::steady_clock::now();
        hipError_t e = hipSuccess;
        if (!admitted.empty()) {
            int max_nodes = 1;
            for (int s : admitted) {
                max_nodes = std::max(max_nodes, graphs[t.h_slots[s].graph].nodes);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(t.d_slots + s, t.h_slots + s, sizeof(BatchSlot), hipMemcpyHostToDevice,
                                       t.stream);
            }
            if (e == hipSuccess)
                e = hipMemcpyAsync(t.d_list, admitted.data(), sizeof(int) * admitted.size(), hipMemcpyHostToDevice,
                                   t.stream);
            if (e == hipSuccess) e = launch_batch_seed(a, t.d_list, (int) admitted.size(), max_nodes, t.stream);
            // the seed list buffer is reused next tick: wait for the seeding
            if (e == hipSuccess) e = hipStreamSynchronize(t.stream);
        }
        // One tick: levels_per_tick levels of every running query.
        if (e == hipSuccess) e = hipGraphLaunch(t.graph_exec, t.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(t.h_slots, t.d_slots, sizeof(BatchSlot) * cfg.max_slots, hipMemcpyDeviceToHost,
                               t.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(t.stream);
        double tms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tick0).count();
        if (e != hipSuccess) {
            st = host_hip_fail(e, "server tick");
            continue;
        }
        // Hand finished queries to the copy thread.  A finished slot stays
        // done (idle) on the device until a new query is seeded into it.
        int handed = 0;
        {
            std::lock_guard<std::mutex> lk(mu);
            stats.ticks++;
            stats.tick_ms += tms;
            for (int s = 0; s < cfg.max_slots; s++) {
                if (slot_ticket[s] < 0 || slot_draining[s]) continue;
                const BatchSlot &bs = t.h_slots[s];
                if (!bs.active || !bs.done) continue;
                const Request &r = reqs[slot_ticket[s]];
                DrainItem it;
                it.slot = s;
                it.ticket = slot_ticket[s];
                it.d_depth = bs.depth;
                it.bytes = sizeof(int) * (size_t) graphs[r.graph].nodes;
                it.out = r.out;
                it.levels = bs.iteration;
                it.status = bs.overflow ? BFS_ERR_OVERFLOW : BFS_OK;
                slot_draining[s] = 1;
                running--;
                draining++;
                drain_q.push_back(it);
                handed++;
            }
        }
        if (handed) cv_drain.notify_one();
    }
}

extern "C" {

bfs_status bfs_server_create(bfs_graph_pool *pool, const bfs_server_config *cfg, bfs_server **out)
{
    if (!pool || !cfg || !out) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    bfs_server_config c = *cfg;
    if (c.max_slots <= 0) c.max_slots = 64;
    if (c.max_slots > kMaxSlots) return host_fail(BFS_ERR_INVALID_ARG, "max_slots must be <= 256");
    if (c.levels_per_tick <= 0) c.levels_per_tick = 2;
    if (c.hbm_budget_bytes <= 0) return host_fail(BFS_ERR_INVALID_ARG, "hbm_budget_bytes must be positive");
    bfs_server *s = new bfs_server();
    s->pool = pool;
    s->cfg = c;
    auto fail_out = [&](bfs_status st) {
        if (s->copy_stream) (void) hipStreamDestroy(s->copy_stream);
        s->t.release();
        if (s->d_arena) (void) hipFree(s->d_arena);
        if (s->d_table) (void) hipFree(s->d_table);
        delete s;
        return st;
    };
    bfs_status st = s->t.init(pool->device, c.max_slots);
    if (st != BFS_OK) return fail_out(st);
    {
        std::lock_guard<std::mutex> lk(pool->mu);
        s->graphs = pool->graphs;
        if (!s->graphs.empty()) {
            if (hipMalloc(&s->d_table, sizeof(BatchGraph) * s->graphs.size()) != hipSuccess ||
                hipMemcpy(s->d_table, pool->d_table, sizeof(BatchGraph) * s->graphs.size(),
                          hipMemcpyDeviceToDevice) != hipSuccess)
                return fail_out(host_fail(BFS_ERR_OOM, "copy graph table"));
        }
    }
    if (s->graphs.empty()) return fail_out(host_fail(BFS_ERR_INVALID_ARG, "pool has no graphs"));
    if (hipMalloc(&s->d_arena, (size_t) c.hbm_budget_bytes) != hipSuccess)
        return fail_out(host_fail(BFS_ERR_OOM, "allocate the query-state budget"));
    s->arena.reset((size_t) c.hbm_budget_bytes);
    if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess)
        return fail_out(host_fail(BFS_ERR_HIP, "create the result-copy stream"));
    s->slot_ticket.assign(c.max_slots, -1);
    s->slot_draining.assign(c.max_slots, 0);
    s->t0 = std::chrono::steady_clock::now();
    s->busy_start = s->t0;
    s->drainer = std::thread([s] { s->drain(); });
    s->thread = std::thread([s] { s->run(); });
    *out = s;
    return BFS_OK;
}

bfs_status bfs_server_submit(bfs_server *s, int32_t graph_id, int32_t src, int32_t *h_depths_out, int64_t *ticket)
{
    if (!s || !ticket) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    if (graph_id < 0 || graph_id >= (int32_t) s->graphs.size()) return host_fail(BFS_ERR_INVALID_ARG, "unknown graph");
    const PoolGraph &g = s->graphs[graph_id];
    if (src < 0 || src >= g.nodes) return host_fail(BFS_ERR_INVALID_ARG, "source out of range");
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.submitted++;
    if (s->fatal != BFS_OK) return s->fatal;
    if ((int64_t) slot_layout(g.nodes).bytes > s->cfg.hbm_budget_bytes) {
        s->stats.rejected++;
        return host_fail(BFS_ERR_REJECTED, "query state exceeds the HBM budget");
    }
    Request r;
    r.ticket = s->next_ticket++;
    r.graph = graph_id;
    r.src = src;
    r.out = h_depths_out;
    r.submit_ms = wall_ms(s->t0);
    s->reqs[r.ticket] = r;
    s->pending.push_back(r.ticket);
    *ticket = r.ticket;
    s->cv_work.notify_one();
    return BFS_OK;
}

bfs_status bfs_server_wait(bfs_server *s, int64_t ticket, int32_t timeout_ms, bfs_request_info *info)
{
    if (!s) return host_fail(BFS_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return host_fail(BFS_ERR_INVALID_ARG, "unknown ticket");
    // A ticket another waiter has already collected counts as finished (no
    // entry is ever re-created here); this waiter then reports it unknown.
    auto ready = [&] {
        auto f = s->reqs.find(ticket);
        return f == s->reqs.end() || f->second.done;
    };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return host_fail(BFS_ERR_TIMEOUT, "wait timed out");
    it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return host_fail(BFS_ERR_INVALID_ARG, "ticket collected by another waiter");
    Request r = it->second;
    s->reqs.erase(it);
    if (info) {
        info->ticket = r.ticket;
        info->graph_id = r.graph;
        info->src = r.src;
        info->levels = r.levels;
        info->status = r.status;
        info->submit_ms = r.submit_ms;
        info->admit_ms = r.admit_ms;
        info->finish_ms = r.finish_ms;
    }
    return r.status;
}

bfs_status bfs_server_get_stats(bfs_server *s, bfs_server_stats *stats)
{
    if (!s || !stats) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *stats = s->stats;
    return BFS_OK;
}

bfs_status bfs_host_alloc_pinned(int64_t bytes, void **ptr)
{
    if (!ptr || bytes <= 0) return host_fail(BFS_ERR_INVALID_ARG, "invalid pinned allocation");
    *ptr = nullptr;
    if (hipHostMalloc(ptr, (size_t) bytes, hipHostMallocDefault) != hipSuccess)
        return host_fail(BFS_ERR_OOM, "pinned host allocation");
    return BFS_OK;
}

void bfs_host_free_pinned(void *ptr)
{
    if (ptr) (void) hipHostFree(ptr);
}

void bfs_server_destroy(bfs_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->thread.joinable()) s->thread.join();
    // the server thread returns once every result is delivered (or on a
    // fatal error); stop the copy thread after whatever it still holds
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->drain_stop = true;
    }
    s->cv_drain.notify_all();
    if (s->drainer.joinable()) s->drainer.join();
    if (s->copy_stream) (void) hipStreamDestroy(s->copy_stream);
    s->t.release();
    if (s->d_arena) (void) hipFree(s->d_arena);
    if (s->d_table) (void) hipFree(s->d_table);
    delete s;
}

}  // extern "C"
