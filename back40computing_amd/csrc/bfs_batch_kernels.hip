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
    unsigned char alive[kMaxSlots];
    int prefix[kMaxSlots + 1];
    int wave_tot[kWaves];
    // visited-edge sums per query: through the previous step (vsum) and
    // through the step being planned (vtop)
    unsigned long long vsum[kMaxSlots];
    unsigned long long vtop[kMaxSlots];
    unsigned char bits[kMaxSlots];  // the current frontier bitmap is exact (BatchSlot::bits)
    int grid_t;              // grid-tier row threshold of this level
    int chunk;               // grid-tier chunk of this level
    int bu_tile;             // kBlock-vertex blocks per bottom-up tile of this level
};

__device__ __forceinline__ unsigned next_item_batch(unsigned *counter, Lds<false> &lds, unsigned k = 1u)
{
    __syncthreads();
    if (threadIdx.x == 0) lds.tile = add_agent(counter, k);
    __syncthreads();
    return lds.tile;
}

__device__ __forceinline__ unsigned long long ld64(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Direction rule of the direction-optimizing search, per query.  ve is the
// visited-edge sum through this step, built from counters that stay fixed
// during the level (the running visited_edges total changes while other
// workgroups already work on the level, and every workgroup must plan the
// same directions).
__device__ __forceinline__ int batch_direction(const BatchArgs &a, const BatchSlot &sl, const BatchGraph &g,
                                               int phase, int step, unsigned n, unsigned long long ve)
{
    if (!a.direction) return kPhaseExpandContract;
    unsigned long long mf = ld64(&sl.fedges[step % 3]);
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
            // the previous level's overflow flag: fixed during this level
            if (ld_agent(&sl.ovf[step % 3])) n = 0;
            L.vtop[tid] = L.vsum[tid] + ld64(&sl.fedges[step % 3]);
            if (n > 0) {
                dir = batch_direction(a, sl, g, L.phase[tid], step, (unsigned) n, L.vtop[tid]);
                // bottom-up: kBlock-vertex blocks here, grouped into tiles below
                tiles = dir == kPhaseBottomUp ? (g.nodes + kBlock - 1) / kBlock : (n + kBlock - 1) / kBlock;
                if (blockIdx.x == 0) {
                    st_agent(&sl.qlen[(step + 2) % 3], 0u);
                    st_agent(&sl.ovf[(step + 2) % 3], 0u);
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
    if (blockIdx.x == 0 && tid == 0) {
        st_agent(&a.ctl->steal[(level + 1) & 1], 0u);
        st_agent(&a.ctl->hub_steal[(level + 1) & 1], 0u);
        __hip_atomic_store(&a.ctl->hub_alloc[(level + 1) & 1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // Bottom-up tile size of this level: up to bu_tile blocks, fewer when
    // the level has too few blocks to give every workgroup about eight
    // tiles (one tile of a small graph would otherwise hold a whole
    // level up while the grid waits at the barrier).
    int bu_blocks = 0;
    (void) block_exclusive_sum(tid < a.nslots && L.dir[tid] == kPhaseBottomUp ? tiles : 0, L.wave_tot, &bu_blocks);
    const int bu_tile = max(1, min(a.bu_tile, bu_blocks / (8 * (int) gridDim.x)));
    if (tid < a.nslots && L.dir[tid] == kPhaseBottomUp) tiles = (tiles + bu_tile - 1) / bu_tile;
    if (tid == 0) L.bu_tile = bu_tile;
    int total = 0;
    int bu_total = 0;
    (void) block_exclusive_sum(tid < a.nslots && L.dir[tid] == kPhaseBottomUp ? tiles : 0, L.wave_tot, &bu_total);
    int excl = block_exclusive_sum(tiles, L.wave_tot, &total);
    if (tid < a.nslots) L.prefix[tid] = excl;
    if (tid == 0) L.prefix[a.nslots] = total;
    // A level of fewer tiles than workgroups leaves most of them idle while
    // a few gather long rows alone: every row long enough for the wave tier
    // then goes to the grid tier, in workgroup-sized chunks.
    const bool narrow = total < (int) gridDim.x;
    L.grid_t = narrow ? min(a.grid_t, max(a.wave_t, kWave)) : a.grid_t;
    L.chunk = narrow ? kBlock : kHubChunk;
    if (blockIdx.x == 0 && tid == 0) a.ctl->hub_chunk[level & 1] = L.chunk;
    if (blockIdx.x == 0 && tid == 0 && level < kLevelLog) {
        a.ctl->level_clock[level] = wall_clock64();
        a.ctl->level_td_tiles[level] = (unsigned) (total - bu_total);
        a.ctl->level_bu_tiles[level] = (unsigned) bu_total;
    }
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

// Per-wave output of the query whose tiles the workgroup is processing: the
// LDS-staged queue and the edge count stay open across consecutive tiles of
// the same query and are flushed when the query changes or the level ends,
// so each query's counters see few atomics however many tiles it has.
struct SlotOut {
    WaveQueue<false> wq;
    unsigned fe;
    int q;
    int step;
};

__device__ __forceinline__ void out_close(SlotOut &o, const BatchArgs &a)
{
    if (o.q < 0) return;
    o.wq.flush();
    slot_edges(a.slots[o.q], o.step, o.fe);
    o.q = -1;
}

__device__ __forceinline__ void out_open(SlotOut &o, const BatchArgs &a, BatchLds &L, int q)
{
    if (o.q == q) return;
    out_close(o, a);
    BatchSlot &sl = a.slots[q];
    const int step = L.step[q];
    o.wq.buf = L.base.wq[threadIdx.x / kWave];
    o.wq.pbuf = nullptr;
    o.wq.count = 0;
    o.wq.out = sl.q[(step + 1) & 1];
    o.wq.pout = nullptr;
    o.wq.out_len = &sl.qlen[(step + 1) % 3];
    o.wq.cap = sl.cap;
    o.wq.overflow = (unsigned *) &sl.overflow;
    o.wq.overflow_step = &sl.ovf[(step + 1) % 3];
    o.fe = 0;
    o.q = q;
    o.step = step;
}

// Top-down edge visit of query q: exact claim, label, frontier edge count,
// append to the query's next frontier.
struct TdEdge {
    BatchSlot &sl;
    const BatchGraph &g;
    int next_depth;
    SlotOut &o;
    bool coherent;
    unsigned *bits;  // next frontier's bitmap to mark claims in, or null
    __device__ __forceinline__ void operator()(bool ok, int u, int par) const
    {
        bool fresh = ok && claim_exact(sl.visited, u, coherent);
        if (fresh) {
            sl.depth[u] = next_depth;
            o.fe += (unsigned) (g.row[u + 1] - g.row[u]);
            if (bits) __hip_atomic_fetch_or(bits + (u >> 5), 1u << (u & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        o.wq.push(fresh, u, par);
    }
};

// Frontier bitmaps across top-down levels.  A bottom-up level reads the
// current frontier from fbm[d & 1] when it is exact (BatchSlot::bits) and
// otherwise tests labels (depth == d), which costs a random label read per
// neighbour.  A top-down level d marks its claims in fbm[(d + 1) & 1] when
// that bitmap is clear -- when level d - 1 was top-down too, since every
// top-down level clears fbm[d & 1], which nothing reads during it -- so
// the usual top-down, top-down, bottom-up sequence reaches its first
// bottom-up level with an exact bitmap.  A bottom-up level writes its whole
// next-frontier bitmap, so it is always exact.
__device__ __forceinline__ unsigned *td_bits(const BatchLds &L, BatchSlot &sl, int q)
{
    return L.phase[q] != kPhaseBottomUp ? sl.fbm[(L.it[q] + 1) & 1] : nullptr;
}

// One top-down (expand-contract) tile of query q.  Rows of at least grid_t
// neighbours are published to the batch-wide grid-tier list (tagged with q).
__device__ void batch_td_tile(const BatchArgs &a, BatchLds &L, int q, unsigned local, SlotOut &o, int level)
{
    BatchSlot &sl = a.slots[q];
    const BatchGraph g = a.graphs[sl.graph];
    const int step = L.step[q];
    const unsigned n = (unsigned) L.n[q];
    HubList hub{a.hubs[level & 1], a.hub_cap, &a.ctl->hub_alloc[level & 1], L.grid_t, q, L.chunk};
    unsigned i = local * kBlock + threadIdx.x;
    int v = -1, beg = 0, end = 0;
    if (i < n) {
        v = ld_queue(sl.q[step & 1] + i);
        beg = g.row[v];
        end = g.row[v + 1];
    }
    {
        // this tile's share of clearing fbm[d & 1] (see td_bits)
        const unsigned words = ((unsigned) g.nodes + 31) >> 5;
        const unsigned ntiles = (n + kBlock - 1) / kBlock;
        const unsigned share = (words + ntiles - 1) / ntiles;
        unsigned *old = sl.fbm[L.it[q] & 1];
        const unsigned w_end = min(words, (local + 1) * share);
        for (unsigned w = local * share + threadIdx.x; w < w_end; w += kBlock) old[w] = 0u;
    }
    gather_rows<false>(beg, end, v, g.col, a.wave_t, a.cta_t, hub, L.base,
                       TdEdge{sl, g, L.it[q] + 1, o, a.coherent_filter != 0, td_bits(L, sl, q)});
}

// Grid-tier phase of a level: every workgroup takes chunks of
// the rows the top-down tiles published.  Follows a grid barrier (or kernel
// boundary) after the tiles.
__device__ void batch_hub_phase(const BatchArgs &a, BatchLds &L, int level)
{
    unsigned long long al = ld64(&a.ctl->hub_alloc[level & 1]);
    unsigned entries = (unsigned) (al >> 32);
    if (entries > a.hub_cap) entries = a.hub_cap;
    if (entries == 0) return;
    const HubEntry *list = a.hubs[level & 1];
    HubEntry last = list[entries - 1];
    // the chunk the level's tiles published with (set when it was planned)
    const int chunk = ld_agent(&a.ctl->hub_chunk[level & 1]);
    unsigned total = last.first_chunk + (unsigned) ((last.end - last.beg + chunk - 1) / chunk);
    SlotOut o;
    o.q = -1;
    // strided grabs of chunks: one counter atomic serves up to grab_max chunks
    const unsigned W = gridDim.x;
    const unsigned grab = max(1u, min((unsigned) a.grab_max, total / (W * 8u)));
    unsigned k_left = 0, ch_next = 0;
    unsigned hint = 0;  // thread 0: entry of this workgroup's previous chunk
    while (true) {
        if (k_left == 0) {
            const unsigned i = next_item_batch(&a.ctl->hub_steal[level & 1], L.base);
            ch_next = (i / W) * W * grab + i % W;
            k_left = grab;
            // first chunks of successive grabs only grow: none is left
            if (ch_next >= total) break;
        } else {
            __syncthreads();  // LDS reuse fence between chunks
        }
        const unsigned ch = ch_next;
        ch_next += W;
        k_left--;
        if (ch >= total) {  // the rest of this grab is past the end
            k_left = 0;
            continue;
        }
        if (threadIdx.x == 0) {
            const unsigned lo = hub_entry_of(list, entries, ch, hint);
            hint = lo;
            HubEntry e = list[lo];
            int b = e.beg + (int) (ch - e.first_chunk) * chunk;
            L.base.cta_beg = b;
            L.base.cta_end = min(e.end, b + chunk);
            L.base.cta_par = e.par;
            L.base.iteration = e.tag;
        }
        __syncthreads();
        const int b = L.base.cta_beg, e = L.base.cta_end, pv = L.base.cta_par, q = L.base.iteration;
        out_open(o, a, L, q);
        BatchSlot &sl = a.slots[q];
        const BatchGraph g = a.graphs[sl.graph];
        TdEdge edge{sl, g, L.it[q] + 1, o, a.coherent_filter != 0, td_bits(L, sl, q)};
        for (int k = b; k < e; k += kBlock) {
            int idx = k + threadIdx.x;
            bool ok = idx < e;
            int u = ok ? __builtin_nontemporal_load(g.col + idx) : -1;
            edge(ok, u, pv);
        }
    }
    out_close(o, a);
}

__device__ __forceinline__ bool batch_hubs_pending(const BatchArgs &a, int level)
{
    return ld64(&a.ctl->hub_alloc[level & 1]) != 0ull;
}

// Bottom-up tile of query q: L.bu_tile consecutive kBlock-vertex blocks, each
// lane owning one vertex of every block (each wave two words of the visited
// and next-frontier bitmaps per block).  The frontier test is the bitmap
// written by a bottom-up previous level, or depth[u] == current depth after
// a top-down one (a vertex labelled during this level holds depth + 1, or
// still -1 if its store is not yet visible, and fails either way).
__device__ __forceinline__ void wave_lds_sync_b()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Queued form of a bottom-up tile (as the single-source step): each wave
// queues the vertices of its slice of the tile's blocks that still search a
// parent and have a non-empty row, and scans them 64 at a time with every
// lane busy; parents found become bits per window of kBatchBuWindow blocks
// (two words per wave and block) written once per word.  The wave's queue
// lives in LDS the batch kernels leave idle during a tile: the history-hash
// area and the scan-tier staging of gather_rows (used only inside a call).
constexpr int kBatchBuWindow = 32;

__device__ void batch_bu_tile(const BatchArgs &a, BatchLds &L, int q, unsigned local, SlotOut &o)
{
    static_assert(kHistory >= kWaves * 256, "per-wave queue of 128 vertices and row starts");
    BatchSlot &sl = a.slots[q];
    const BatchGraph g = a.graphs[sl.graph];
    const int d = L.it[q];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const bool use_bitmap = L.bits[q] != 0;
    const bool symmetric = g.rrow == g.row;
    const unsigned *cur = sl.fbm[d & 1];
    unsigned *nxt = sl.fbm[(d + 1) & 1];
    const unsigned words = ((unsigned) g.nodes + 31) >> 5;
    const unsigned nblocks = (unsigned) (g.nodes + kBlock - 1) / kBlock;
    const unsigned first = local * (unsigned) L.bu_tile;
    const unsigned last = min(nblocks, first + (unsigned) L.bu_tile);
    int *qv = L.base.hist + wave * 256;  // [128] vertices
    int *qb = qv + 128;                  // [128] row starts
    int *qe0 = L.base.sc_excl[wave];     // row ends, entries 0..63
    int *qe1 = L.base.sc_beg[wave];      // row ends, entries 64..127
    unsigned *fb = reinterpret_cast<unsigned *>(L.base.sc_par[wave]);  // found bits of the window

    auto in_frontier = [&](int u) {
        if (use_bitmap) return ((cur[u >> 5] >> (u & 31)) & 1u) != 0;
        return ((sl.visited[u >> 5] >> (u & 31)) & 1u) != 0 && sl.depth[u] == d;
    };
    // Scans the first cnt (1..64) queued vertices of the window starting at block wb0.
    auto scan = [&](int cnt, unsigned wb0) {
        int v = -1, beg = 0, end = 0;
        if (lane < cnt) {
            v = qv[lane];
            beg = qb[lane];
            end = qe0[lane];
        }
        int pos = beg;
        const int lim = end - beg < 64 ? end : min(end, beg + 4);
        int parent = -1;
        while (parent < 0 && pos < lim) {
            int u[4];
            unsigned fw[4];
            int lab[4];
#pragma unroll
            for (int j = 0; j < 4; j++) u[j] = pos + j < lim ? g.rcol[pos + j] : -1;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                fw[j] = u[j] >= 0 ? (use_bitmap ? cur[u[j] >> 5] : sl.visited[u[j] >> 5]) : 0u;
                lab[j] = u[j] >= 0 && !use_bitmap ? sl.depth[u[j]] : d;
            }
#pragma unroll
            for (int j = 3; j >= 0; j--)  // the first hit in row order wins
                if (u[j] >= 0 && ((fw[j] >> (u[j] & 31)) & 1u) && lab[j] == d) parent = u[j];
            pos += 4;
        }
        uint64_t lm = ballot(parent < 0 && lim < end);
        while (lm) {
            const int l = uniform(__builtin_ctzll(lm));
            lm &= lm - 1;
            const int b = readlane(lim, l), e = readlane(end, l);
            int hit = -1;
            for (int r = b; r < e; r += kWave) {
                const int idx = r + lane;
                const int uu = idx < e ? g.rcol[idx] : -1;
                const uint64_t hm = ballot(uu >= 0 && in_frontier(uu));
                if (hm) {
                    hit = readlane(uu, uniform(__builtin_ctzll(hm)));
                    break;
                }
            }
            if (lane == l) parent = hit;
        }
        const bool got = parent >= 0;
        if (got) {
            sl.depth[v] = d + 1;
            // out-degree: the scanned in-row's length when the graph is symmetric
            o.fe += (unsigned) (symmetric ? end - beg : g.row[v + 1] - g.row[v]);
            const unsigned wl = 2u * (((unsigned) v / kBlock) - wb0) + (((unsigned) v >> 5) & 1u);
            __hip_atomic_fetch_or(&fb[wl], 1u << (v & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
        o.wq.push(got, v, parent);
    };
    // Moves queue entries [64, 64 + n) (n < 64) to the front.
    auto shift = [&](int n) {
        int v = 0, b = 0, e = 0;
        if (lane < n) {
            v = qv[kWave + lane];
            b = qb[kWave + lane];
            e = qe1[lane];
        }
        wave_lds_sync_b();
        if (lane < n) {
            qv[lane] = v;
            qb[lane] = b;
            qe0[lane] = e;
        }
        wave_lds_sync_b();
    };

    // software pipeline: the next block's visited word and row bounds are
    // in flight while the current one is queued and scanned
    unsigned vis_n = ~0u;
    int beg_n = 0, end_n = 0;
    auto load_block = [&](unsigned b) {
        const int v = (int) (b * kBlock + threadIdx.x);
        const bool ok = b < last && v < g.nodes;
        vis_n = ok ? sl.visited[(unsigned) v >> 5] : ~0u;
        beg_n = ok ? g.rrow[v] : 0;
        end_n = ok ? g.rrow[v + 1] : 0;
    };
    if (first < last) load_block(first);
    for (unsigned wb0 = first; wb0 < last; wb0 += kBatchBuWindow) {
        const unsigned wb_end = min(last, wb0 + kBatchBuWindow);
        fb[lane] = 0u;  // 64 lanes: two words per block of the window
        wave_lds_sync_b();
        // one loop with a single scan site (inlining the scan twice costs
        // registers, and with them workgroups per CU): queue the next block
        // while fewer than 64 wait and blocks remain, scan otherwise
        int qn = 0;
        unsigned b = wb0;
        while (true) {
            if (qn >= kWave || (b >= wb_end && qn > 0)) {
                const int cnt = min(qn, kWave);
                wave_lds_sync_b();
                scan(cnt, wb0);
                if (qn > cnt) shift(qn - cnt);
                qn -= cnt;
                continue;
            }
            if (b >= wb_end) break;
            const unsigned vis = vis_n;
            const int beg = beg_n, end = end_n;
            if (b + 1 < last) load_block(b + 1);
            const int v = (int) (b * kBlock + threadIdx.x);
            const bool todo = v < g.nodes && ((vis >> (v & 31)) & 1u) == 0 && end > beg;
            const uint64_t m = ballot(todo);
            if (todo) {
                const int p = qn + lanes_below(m);
                qv[p] = v;
                qb[p] = beg;
                if (p < kWave) qe0[p] = end;
                else qe1[p - kWave] = end;
            }
            qn += __popcll(m);
            b++;
        }
        wave_lds_sync_b();
        // the window's visited and next-frontier words of this wave's slices
        if (lane < 2 * (int) (wb_end - wb0)) {
            const unsigned w = (wb0 + (unsigned) lane / 2u) * (kBlock / 32) + 2u * (unsigned) wave + ((unsigned) lane & 1u);
            if (w < words) {
                const unsigned bits = fb[lane];
                if (bits) sl.visited[w] |= bits;
                nxt[w] = bits;  // whole word: also clears it
            }
        }
        wave_lds_sync_b();
    }
}

// Runs the tiles of one planned level.
__device__ void batch_tiles(const BatchArgs &a, BatchLds &L, int level, int total)
{
    SlotOut o;
    o.q = -1;
    // Grabs: one counter atomic hands a workgroup `grab` tiles, so the
    // single counter address (whose device-scope atomics serialise) is not
    // the bottleneck of levels with many tiles.  Strided: counter value i
    // takes tiles i0, i0 + W, ... (W = gridDim.x), so W consecutive grabs
    // cover W * grab consecutive tiles and the tiles in flight across the
    // grid stay a window about W wide.  Consecutive: counter value t0 takes
    // [t0, t0 + grab).
    const unsigned W = gridDim.x;
    const unsigned grab = (unsigned) max(1, min(a.grab_max, total / ((int) W * 8)));
    const unsigned step = a.strided ? W : 1u;
    while (true) {
        unsigned t0;
        if (a.strided) {
            const unsigned i = next_item_batch(&a.ctl->steal[level & 1], L.base);
            t0 = (i / W) * W * grab + i % W;
        } else {
            t0 = next_item_batch(&a.ctl->steal[level & 1], L.base, grab);
        }
        if (t0 >= (unsigned) total) break;
        for (unsigned t = t0, k = 0; k < grab && t < (unsigned) total; k++, t += step) {
            // last query whose first tile is <= t (queries without tiles
            // repeat the prefix value and are skipped by taking the last match)
            int lo = 0, hi = a.nslots - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if ((unsigned) L.prefix[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            unsigned local = t - (unsigned) L.prefix[lo];
            out_open(o, a, L, lo);
            if (L.dir[lo] == kPhaseBottomUp) batch_bu_tile(a, L, lo, local, o);
            else batch_td_tile(a, L, lo, local, o, level);
        }
    }
    out_close(o, a);
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
        L.vsum[tid] = ld64(&sl.vbase);
        L.bits[tid] = ld_agent(&sl.bits);
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
        if (batch_hubs_pending(a, level)) {
            batch_hub_phase(a, L, level);
            grid_barrier(&a.ctl->bar);
        }
        // every workgroup advances its own copy of the cursors identically
        int tid = threadIdx.x;
        if (tid < a.nslots && L.n[tid] > 0) {
            L.step[tid]++;
            L.it[tid]++;
            L.bits[tid] = L.dir[tid] == kPhaseBottomUp || L.phase[tid] != kPhaseBottomUp;
            L.phase[tid] = L.dir[tid];
            L.vsum[tid] = L.vtop[tid];
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
            sl.vbase = L.vsum[tid];
            sl.bits = L.bits[tid];
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

// Grid-tier chunks of a level run as their own launch.
__global__ __launch_bounds__(kBlock) void batch_hub_kernel(BatchArgs a)
{
    __shared__ BatchLds L;
    int level = ld_agent(&a.ctl->level);
    if (!batch_hubs_pending(a, level)) return;
    batch_load_cursors(a, L);
    batch_hub_phase(a, L, level);
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
            sl.bits = sl.next_phase == kPhaseBottomUp || sl.phase != kPhaseBottomUp;
            sl.phase = sl.next_phase;
            sl.vbase += sl.fedges[sl.step % 3];
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

// Resets depth, visited and frontier bitmaps of the queries listed in `list` (blockIdx.y picks
// the query).
__global__ void batch_reset_kernel(BatchArgs a, const int *list)
{
    BatchSlot &sl = a.slots[list[blockIdx.y]];
    const int nodes = a.graphs[sl.graph].nodes;
    const int words = (nodes + 31) >> 5;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += gridDim.x * blockDim.x) {
        sl.depth[i] = -1;
        if (i < words) {
            sl.visited[i] = 0u;
            sl.fbm[0][i] = 0u;  // top-down levels mark claims in clear bitmaps
            sl.fbm[1][i] = 0u;
        }
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
    sl.ovf[0] = sl.ovf[1] = sl.ovf[2] = 0;
    sl.vbase = 0;
    sl.bits = 0;  // fbm[0] does not hold the source: a bottom-up level 0 tests labels
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
    return launch_coresident(reinterpret_cast<const void *>(&batch_persistent_kernel), dim3(grid), dim3(kBlock), args,
                             st);
}

hipError_t launch_batch_level(const BatchArgs &a, int grid, int *alive_out, hipStream_t st)
{
    hipLaunchKernelGGL(batch_level_kernel, dim3(grid), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(batch_hub_kernel, dim3(grid), dim3(kBlock), 0, st, a);
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
