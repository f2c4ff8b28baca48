// bfs_batch_kernels.hip -- batched BFS kernels for gfx950.
//
// Many single-source queries over a pool of device-resident graphs are
// advanced one level at a time together: at each level every workgroup
// plans the level for all queries (frontier length, direction, tile count),
// then workgroups take tiles from one batch-wide counter; a tile belongs to
// one query and is either a top-down tile (256 frontier vertices, gathered
// by the CTA / wave / scan tiers with exact claims) or a bottom-up tile (256
// consecutive vertices looking for a parent at the current depth).  Levels
// are separated by the grid barrier of a persistent launch, or by kernel
// boundaries when each level is a launch (those launches are captured into a
// hipGraph by the host).  The reference has no batched mode; the per-query
// step semantics are those of the single-source kernels.
#include "bfs_batch.hpp"
#include "bfs_batch_internal.hpp"
#include "bfs_internal.hpp"

namespace b40c_hip {

struct BatchLds {
    Lds<false> base;
    int n[kMaxSlots];        // frontier length of each query this level
    int dir[kMaxSlots];      // direction chosen for this level
    int step[kMaxSlots];     // cursor copies
    int it[kMaxSlots];
    int phase[kMaxSlots];
    int alive[kMaxSlots];
    int prefix[kMaxSlots + 1];
    int wave_tot[kWaves];
};

__device__ __forceinline__ unsigned next_item_batch(unsigned *counter, Lds<false> &lds)
{
    __syncthreads();
    if (threadIdx.x == 0) lds.tile = add_agent(counter, 1u);
    __syncthreads();
    return lds.tile;
}

__device__ __forceinline__ unsigned long long ld64(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix sum over the workgroup (one value per thread): wave DPP
// sums, then the four wave totals through LDS.
__device__ __forceinline__ int block_exclusive_sum(int x, int *wave_tot, int *total)
{
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    int incl = wave_inclusive_sum(x);
    if (lane == kWave - 1) wave_tot[wave] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < kWaves; w++) {
        int t = wave_tot[w];
        if (w < wave) off += t;
        tot += t;
    }
    *total = tot;
    __syncthreads();
    return off + incl - x;
}

// Direction rule of the direction-optimizing search, per query.
__device__ __forceinline__ int batch_direction(const BatchArgs &a, const BatchSlot &sl, const BatchGraph &g,
                                               int phase, int step, unsigned n)
{
    if (!a.direction) return kPhaseExpandContract;
    unsigned long long mf = ld64(&sl.fedges[step % 3]);
    unsigned long long ve = ld64(&sl.visited_edges);
    unsigned long long mu = g.edges > ve ? g.edges - ve : 0ull;
    if (phase == kPhaseExpandContract) {
        if (mf * (unsigned long long) a.do_alpha > mu) return kPhaseBottomUp;
        return kPhaseExpandContract;
    }
    if ((unsigned long long) n * (unsigned long long) a.do_beta < (unsigned long long) g.nodes)
        return kPhaseExpandContract;
    return kPhaseBottomUp;
}

// Plans level `level` for every query from the cursor copies in LDS:
// frontier lengths, directions and the tile prefix.  Block 0 also zeroes the
// counters the next level appends to and marks finished queries.  Returns
// the total number of tiles (grid-uniform).
__device__ int batch_plan(const BatchArgs &a, BatchLds &L, int level)
{
    const int tid = threadIdx.x;
    int tiles = 0;
    if (tid < a.nslots) {
        BatchSlot &sl = a.slots[tid];
        int n = 0, dir = L.phase[tid];
        if (L.alive[tid]) {
            const BatchGraph &g = a.graphs[sl.graph];
            int step = L.step[tid];
            unsigned len = ld_agent(&sl.qlen[step % 3]);
            n = (int) (len < sl.cap ? len : sl.cap);
            if (ld_agent((const unsigned *) &sl.overflow)) n = 0;
            if (n > 0) {
                dir = batch_direction(a, sl, g, L.phase[tid], step, (unsigned) n);
                tiles = dir == kPhaseBottomUp ? (g.nodes + kBlock - 1) / kBlock : (n + kBlock - 1) / kBlock;
                if (blockIdx.x == 0) {
                    st_agent(&sl.qlen[(step + 2) % 3], 0u);
                    __hip_atomic_store(&sl.fedges[(step + 2) % 3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                L.alive[tid] = 0;
                if (blockIdx.x == 0) {
                    sl.done = 1;
                    sl.t_end = wall_clock64();
                }
            }
        }
        L.n[tid] = n;
        L.dir[tid] = dir;
        if (blockIdx.x == 0) {
            sl.next_phase = dir;
            sl.ran = n > 0;
        }
    }
    if (blockIdx.x == 0 && tid == 0) st_agent(&a.ctl->steal[(level + 1) & 1], 0u);
    int total = 0;
    int excl = block_exclusive_sum(tiles, L.wave_tot, &total);
    if (tid < a.nslots) L.prefix[tid] = excl;
    if (tid == 0) L.prefix[a.nslots] = total;
    __syncthreads();
    return total;
}

__device__ __forceinline__ void slot_edges(BatchSlot &sl, int step, unsigned fe)
{
    int lo = wave_inclusive_sum((int) (fe & 0xFFFFu));
    int hi = wave_inclusive_sum((int) (fe >> 16));
    unsigned long long tot = ((unsigned long long) (unsigned) readlane(hi, kWave - 1) << 16) +
                             (unsigned long long) (unsigned) readlane(lo, kWave - 1);
    if (tot != 0 && lane_id() == 0) {
        __hip_atomic_fetch_add(&sl.fedges[(step + 1) % 3], tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&sl.visited_edges, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One top-down (expand-contract) tile of query q.
__device__ void batch_td_tile(const BatchArgs &a, BatchLds &L, int q, unsigned local)
{
    BatchSlot &sl = a.slots[q];
    const BatchGraph g = a.graphs[sl.graph];
    const int step = L.step[q], next_depth = L.it[q] + 1;
    const unsigned n = (unsigned) L.n[q];
    const int wave = threadIdx.x / kWave;
    WaveQueue<false> wq;
    wq.buf = L.base.wq[wave];
    wq.pbuf = nullptr;
    wq.count = 0;
    wq.out = sl.q[(step + 1) & 1];
    wq.pout = nullptr;
    wq.out_len = &sl.qlen[(step + 1) % 3];
    wq.cap = sl.cap;
    wq.overflow = (unsigned *) &sl.overflow;
    HubList nohub{nullptr, 0, nullptr, 0x7fffffff};
    unsigned fe = 0;
    unsigned i = local * kBlock + threadIdx.x;
    int v = -1, beg = 0, end = 0;
    if (i < n) {
        v = sl.q[step & 1][i];
        beg = g.row[v];
        end = g.row[v + 1];
    }
    gather_rows<false>(beg, end, v, g.col, a.wave_t, a.cta_t, nohub, L.base, [&](bool ok, int u, int par) {
        bool fresh = ok && claim_exact(sl.visited, u);
        if (fresh) {
            sl.depth[u] = next_depth;
            fe += (unsigned) (g.row[u + 1] - g.row[u]);
        }
        wq.push(fresh, u, par);
    });
    wq.flush();
    slot_edges(sl, step, fe);
}

// One bottom-up tile of query q: vertices [local*256, local*256+256).  The
// frontier test is depth[u] == current depth (a vertex labelled during this
// level holds depth + 1, so it never passes); each wave owns two bitmap words.
__device__ void batch_bu_tile(const BatchArgs &a, BatchLds &L, int q, unsigned local)
{
    BatchSlot &sl = a.slots[q];
    const BatchGraph g = a.graphs[sl.graph];
    const int step = L.step[q], d = L.it[q];
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    WaveQueue<false> wq;
    wq.buf = L.base.wq[wave];
    wq.pbuf = nullptr;
    wq.count = 0;
    wq.out = sl.q[(step + 1) & 1];
    wq.pout = nullptr;
    wq.out_len = &sl.qlen[(step + 1) % 3];
    wq.cap = sl.cap;
    wq.overflow = (unsigned *) &sl.overflow;
    unsigned fe = 0;
    int v = (int) (local * kBlock + threadIdx.x);
    unsigned w = (unsigned) v >> 5;
    bool valid = v < g.nodes;
    unsigned vis = valid ? sl.visited[w] : ~0u;
    bool todo = valid && ((vis >> (v & 31)) & 1u) == 0;
    int beg = 0, end = 0;
    if (todo) {
        beg = g.rrow[v];
        end = g.rrow[v + 1];
    }
    int len = end - beg, parent = -1;
    if (todo && len < 64) {
        for (int e = beg; e < end; e += 4) {
            int u0 = g.rcol[e];
            int u1 = e + 1 < end ? g.rcol[e + 1] : -1;
            int u2 = e + 2 < end ? g.rcol[e + 2] : -1;
            int u3 = e + 3 < end ? g.rcol[e + 3] : -1;
            if (sl.depth[u0] == d) { parent = u0; break; }
            if (u1 >= 0 && sl.depth[u1] == d) { parent = u1; break; }
            if (u2 >= 0 && sl.depth[u2] == d) { parent = u2; break; }
            if (u3 >= 0 && sl.depth[u3] == d) { parent = u3; break; }
        }
    }
    uint64_t lm = ballot(todo && len >= 64);
    while (lm) {
        int l = uniform(__builtin_ctzll(lm));
        lm &= lm - 1;
        int b = readlane(beg, l), e = readlane(end, l);
        int found = -1;
        for (int k = b; k < e; k += kWave) {
            int idx = k + lane;
            int u = idx < e ? g.rcol[idx] : -1;
            uint64_t hm = ballot(u >= 0 && sl.depth[u] == d);
            if (hm) {
                found = readlane(u, uniform(__builtin_ctzll(hm)));
                break;
            }
        }
        if (lane == l) parent = found;
    }
    bool got = parent >= 0;
    if (got) {
        sl.depth[v] = d + 1;
        fe += (unsigned) (g.row[v + 1] - g.row[v]);
    }
    uint64_t gm = ballot(got);
    if ((lane & 31) == 0 && valid) {
        unsigned bits = (unsigned) (lane == 0 ? gm : gm >> 32);
        if (bits) sl.visited[w] = vis | bits;
    }
    wq.push(got, v, parent);
    wq.flush();
    slot_edges(sl, step, fe);
}

// Runs the tiles of one planned level.
__device__ void batch_tiles(const BatchArgs &a, BatchLds &L, int level, int total)
{
    while (true) {
        unsigned t = next_item_batch(&a.ctl->steal[level & 1], L.base);
        if (t >= (unsigned) total) break;
        // last query whose first tile is <= t (queries without tiles repeat
        // the prefix value and are skipped by taking the last match)
        int lo = 0, hi = a.nslots - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if ((unsigned) L.prefix[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        unsigned local = t - (unsigned) L.prefix[lo];
        if (L.dir[lo] == kPhaseBottomUp) batch_bu_tile(a, L, lo, local);
        else batch_td_tile(a, L, lo, local);
    }
}

__device__ void batch_load_cursors(const BatchArgs &a, BatchLds &L)
{
    for (int i = threadIdx.x; i < kHistory; i += kBlock) L.base.hist[i] = -1;
    if (threadIdx.x == 0) L.base.cta_owner = -1;
    int tid = threadIdx.x;
    if (tid < a.nslots) {
        const BatchSlot &sl = a.slots[tid];
        L.step[tid] = ld_agent(&sl.step);
        L.it[tid] = ld_agent(&sl.iteration);
        L.phase[tid] = ld_agent(&sl.phase);
        L.alive[tid] = ld_agent(&sl.active) && !ld_agent(&sl.done);
    }
    __syncthreads();
}

// All levels of the batch in one cooperative launch.
__global__ __launch_bounds__(kBlock) void batch_persistent_kernel(BatchArgs a)
{
    __shared__ BatchLds L;
    batch_load_cursors(a, L);
    int level = ld_agent(&a.ctl->level);
    while (true) {
        int total = batch_plan(a, L, level);
        if (total == 0) break;
        batch_tiles(a, L, level, total);
        grid_barrier(&a.ctl->bar);
        // every workgroup advances its own copy of the cursors identically
        int tid = threadIdx.x;
        if (tid < a.nslots && L.n[tid] > 0) {
            L.step[tid]++;
            L.it[tid]++;
            L.phase[tid] = L.dir[tid];
        }
        __syncthreads();
        level++;
    }
    if (blockIdx.x == 0) {
        int tid = threadIdx.x;
        if (tid < a.nslots) {
            BatchSlot &sl = a.slots[tid];
            sl.step = L.step[tid];
            sl.iteration = L.it[tid];
            sl.phase = L.phase[tid];
        }
        if (tid == 0) a.ctl->level = level;
    }
}

// One level per launch; batch_advance_kernel follows each level.
__global__ __launch_bounds__(kBlock) void batch_level_kernel(BatchArgs a)
{
    __shared__ BatchLds L;
    batch_load_cursors(a, L);
    int level = ld_agent(&a.ctl->level);
    int total = batch_plan(a, L, level);
    if (total == 0) return;
    batch_tiles(a, L, level, total);
}

// Single workgroup: moves every query that had a frontier one step on and
// counts the queries still running.
__global__ __launch_bounds__(kBlock) void batch_advance_kernel(BatchArgs a, int *alive_out)
{
    int tid = threadIdx.x;
    int alive = 0;
    if (tid < a.nslots) {
        BatchSlot &sl = a.slots[tid];
        if (sl.ran) {
            sl.phase = sl.next_phase;
            sl.step++;
            sl.iteration++;
            sl.ran = 0;
        }
        alive = sl.active && !sl.done;
    }
    __shared__ int wave_tot[kWaves];
    int total = 0;
    (void) block_exclusive_sum(alive, wave_tot, &total);
    if (tid == 0) {
        a.ctl->level++;
        if (alive_out) *alive_out = total;
    }
}

// Resets depth / visited of the queries listed in `list` (blockIdx.y picks
// the query).
__global__ void batch_reset_kernel(BatchArgs a, const int *list)
{
    BatchSlot &sl = a.slots[list[blockIdx.y]];
    const int nodes = a.graphs[sl.graph].nodes;
    const int words = (nodes + 31) >> 5;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += gridDim.x * blockDim.x) {
        sl.depth[i] = -1;
        if (i < words) sl.visited[i] = 0u;
    }
}

// Seeds the listed queries: source labelled 0, queue 0 = {source}.
__global__ void batch_seed_kernel(BatchArgs a, const int *list, int count)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    BatchSlot &sl = a.slots[list[k]];
    const BatchGraph &g = a.graphs[sl.graph];
    int src = sl.src;
    sl.depth[src] = 0;
    sl.visited[src >> 5] |= 1u << (src & 31);
    sl.q[0][0] = src;
    sl.qlen[0] = 1;
    sl.qlen[1] = 0;
    sl.qlen[2] = 0;
    unsigned long long deg = (unsigned long long) (g.row[src + 1] - g.row[src]);
    sl.fedges[0] = deg;
    sl.fedges[1] = 0;
    sl.fedges[2] = 0;
    sl.visited_edges = deg;
    sl.step = 0;
    sl.iteration = 0;
    sl.phase = kPhaseExpandContract;
    sl.overflow = 0;
    sl.ran = 0;
    sl.done = 0;
    sl.t_end = 0;
    sl.t_begin = wall_clock64();
    __threadfence();
    sl.active = 1;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
hipError_t launch_batch_persistent(const BatchArgs &a, int grid, hipStream_t st)
{
    BatchArgs aa = a;
    void *args[] = {&aa};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&batch_persistent_kernel), dim3(grid),
                                      dim3(kBlock), args, 0, st);
}

hipError_t launch_batch_level(const BatchArgs &a, int grid, int *alive_out, hipStream_t st)
{
    hipLaunchKernelGGL(batch_level_kernel, dim3(grid), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(kBlock), 0, st, a, alive_out);
    return hipGetLastError();
}

hipError_t launch_batch_seed(const BatchArgs &a, const int *d_list, int count, int max_nodes, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    int gx = (max_nodes + 255) / 256;
    if (gx > 256) gx = 256;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(gx, count), dim3(256), 0, st, a, d_list);
    hipLaunchKernelGGL(batch_seed_kernel, dim3((count + 255) / 256), dim3(256), 0, st, a, d_list, count);
    return hipGetLastError();
}

int batch_persistent_blocks_per_cu()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void *>(&batch_persistent_kernel),
                                                     kBlock, 0) != hipSuccess)
        return 0;
    return blocks;
}

int batch_level_blocks_per_cu()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void *>(&batch_level_kernel),
                                                     kBlock, 0) != hipSuccess)
        return 0;
    return blocks;
}

}  // namespace b40c_hip
