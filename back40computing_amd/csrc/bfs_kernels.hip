// bfs_kernels.hip -- gfx950 breadth-first-search kernels.
//
// One step of the level loop maps a frontier queue to the next one.  The
// step bodies are:
//   expand-contract  vertex frontier -> gather neighbours -> claim -> vertex
//                    frontier (b40c/graph/bfs/expand_contract_atomic/cta.cuh:665-788)
//   contract-expand  edge frontier -> claim -> gather neighbours -> edge
//                    frontier (b40c/graph/bfs/contract_expand_atomic/cta.cuh:573-709)
//   contract         edge frontier -> claim -> vertex frontier
//                    (b40c/graph/bfs/two_phase/contract_atomic/cta.cuh:494-572)
//   expand           vertex frontier -> gather -> edge frontier
//                    (b40c/graph/bfs/two_phase/expand_atomic/cta.cuh:606-719)
//   filter           edge frontier -> visited-bitmask cull -> edge frontier
//                    (b40c/graph/bfs/two_phase/filter_atomic/cta.cuh:308-388)
// Gathering steps have two sub-phases: the tile phase (frontier tiles taken
// dynamically by workgroups, rows gathered by the CTA / wave / scan tiers)
// and, when some rows were long enough for the grid tier, the hub phase in
// which every workgroup takes fixed-size chunks of those rows.
// Each step runs either inside a persistent, cooperatively launched loop with
// in-kernel grid barriers (the reference's KernelGlobalBarrier kernels) or
// as kernel launches per step.
#include "b40c_primitives.hpp"
#include "bfs_device.hpp"
#include "bfs_internal.hpp"

#include <algorithm>
#include <atomic>

namespace b40c_hip {

template <bool PRED>
__device__ __forceinline__ void init_lds(Lds<PRED> &lds)
{
    for (int i = threadIdx.x; i < kHistory; i += kBlock) lds.hist[i] = -1;
    if (threadIdx.x == 0) {
        lds.cta_owner = -1;
        lds.solo = 0;
    }
    __syncthreads();
}

// Dynamic work assignment: one agent-scope atomic per tile (or chunk).
template <bool PRED>
__device__ __forceinline__ unsigned next_item(unsigned *counter, Lds<PRED> &lds, unsigned k = 1u)
{
    __syncthreads();
    if (threadIdx.x == 0) lds.tile = add_agent(counter, k);
    __syncthreads();
    return lds.tile;
}

// Guided grabs of kBlock-item tiles: one counter atomic hands a workgroup
// `grab` consecutive tiles, about eight grabs per workgroup when a phase has
// many tiles, so the single counter address (whose device-scope atomics
// serialise) stays off the critical path of large levels.
// A phase of at most one tile per workgroup that every workgroup runs needs
// no counter: workgroup b takes tile b (grab 0), and the two round trips of
// the counter atomics (the grab, and the one that finds nothing left) leave
// the level's critical path.  A workgroup running a step alone (lds.solo)
// grabs every tile itself.
struct TileGrab {
    unsigned next, end, grab, tiles;
};

// tiles per counter atomic of a dynamic assignment
__device__ __forceinline__ unsigned tile_grab_size(const SearchArgs &s, unsigned tiles)
{
    return max(1u, min((unsigned) s.grab_max, tiles / (gridDim.x * 8u)));
}

template <bool PRED>
__device__ __forceinline__ TileGrab tile_grab(const SearchArgs &s, unsigned items, const Lds<PRED> &lds)
{
    unsigned tiles = (items + kBlock - 1) / kBlock;
    if (tiles <= gridDim.x && !lds.solo)
        return TileGrab{blockIdx.x, blockIdx.x < tiles ? blockIdx.x + 1 : blockIdx.x, 0u, tiles};
    return TileGrab{0u, 0u, tile_grab_size(s, tiles), tiles};
}

// Next tile index of the workgroup (workgroup-uniform); with a static
// assignment, the phase's tile count once the workgroup's tile is done.
template <bool PRED>
__device__ __forceinline__ unsigned next_tile(unsigned *counter, Lds<PRED> &lds, TileGrab &tg)
{
    if (tg.grab == 0) {
        __syncthreads();  // the LDS reuse fence of a grab
        return tg.next == tg.end ? tg.tiles : tg.next++;
    }
    if (tg.next == tg.end) {
        tg.next = next_item(counter, lds, tg.grab);
        tg.end = tg.next + tg.grab;
    } else {
        __syncthreads();  // same LDS reuse fence as a fresh grab
    }
    return tg.next++;
}

// LDS history hash: true if u may be new to this workgroup.  A hit means
// another lane of this workgroup already took u through the claim, so this
// copy can be dropped without touching global memory.  Racing lanes may both
// miss; the claim that follows settles them.  Cf. HistoryCull / WarpCull
// (b40c/graph/bfs/two_phase/contract_atomic/cta.cuh:276-327).
template <bool PRED>
__device__ __forceinline__ bool history_pass(Lds<PRED> &lds, int u)
{
    unsigned h = ((unsigned) u * 0x9E3779B1u) >> (32 - 10);
    static_assert(kHistory == 1024, "history hash uses 10 bits");
    if (lds.hist[h] == u) return false;
    lds.hist[h] = u;
    return true;
}

template <bool PRED>
__device__ __forceinline__ WaveQueue<PRED> make_queue(const SearchArgs &s, int step, Lds<PRED> &lds)
{
    const int wave = threadIdx.x / kWave;
    WaveQueue<PRED> wq;
    wq.buf = lds.wq[wave];
    if constexpr (PRED) wq.pbuf = lds.wqp.buf[wave];
    else wq.pbuf = nullptr;
    wq.count = 0;
    wq.out = s.q[(step + 1) & 1];
    wq.pout = PRED ? s.pq[(step + 1) & 1] : nullptr;
    wq.out_len = &s.ctl->qlen[(step + 1) % 3];
    wq.cap = s.cap;
    wq.overflow = &s.ctl->overflow;
    wq.overflow_step = &s.ctl->ovf[(step + 1) % 3];
    return wq;
}

// Steps with fewer tiles than the grid has workgroups (a level of a handful
// of hub vertices): most workgroups would idle while a few stride long rows
// alone, so every row the workgroup tier would take goes to the grid tier
// instead, in kBlock-edge chunks.  Grid-uniform (n is read after a barrier).
__device__ __forceinline__ bool narrow_step(unsigned n) { return (n + kBlock - 1) / kBlock < gridDim.x; }

// Grid-tier chunk of a step over n items.  Marked steps (see
// body_expand_mark) keep kHubChunk chunks even when narrow: their chunks
// are gathered four workgroup-widths at a time.
__device__ __forceinline__ int hub_chunk(unsigned n, bool marked = false)
{
    return narrow_step(n) && !marked ? kBlock : kHubChunk;
}

// Grid-tier list of a gathering step over n input items.
__device__ __forceinline__ HubList hub_list(const SearchArgs &s, int step, bool enabled, unsigned n,
                                            bool marked = false)
{
    HubList h;
    h.entries = s.hubs[step & 1];
    h.cap = s.hub_cap;
    h.alloc = &s.ctl->hub_alloc[step & 1];
    h.grid_t = enabled ? s.grid_t : 0x7fffffff;
    // a step of fewer tiles than workgroups leaves most workgroups idle:
    // every row long enough for the wave tier goes to the grid instead
    if (enabled && narrow_step(n)) h.grid_t = min(h.grid_t, max(s.wave_t, kWave));
    h.tag = 0;
    h.chunk = hub_chunk(n, marked);
    return h;
}

// End of a step for the whole workgroup (every thread calls it): the waves'
// staged queue entries are reserved with one atomic, and their degree sums
// and depth note are folded into one update each.  Per-wave updates would
// put four times as many device-scope atomics on the same few addresses,
// where they serialise across the grid.
//
// `uncounted` (per lane) adds frontier members that are not staged in the
// queue: a bottom-up level records its output only in the frontier bitmap,
// and the queue length alone carries its size.
template <bool PRED>
__device__ void finish_step(WaveQueue<PRED> &wq, Lds<PRED> &lds, Control *ctl, int step, bool labeled, int depth,
                            unsigned fe, bool track, int uncounted = 0)
{
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    int lo = wave_inclusive_sum((int) (fe & 0xFFFFu));
    int hi = wave_inclusive_sum((int) (fe >> 16));
    int extra = wave_inclusive_sum(uncounted);
    bool any = ballot(labeled) != 0;
    if (lane == kWave - 1) {  // the inclusive sums end in the last lane
        lds.ep_count[wave] = (unsigned) wq.count;
        lds.ep_extra[wave] = (unsigned) extra;
        lds.ep_fe[wave] = ((unsigned long long) (unsigned) hi << 16) + (unsigned long long) (unsigned) lo;
        lds.ep_lab[wave] = any ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned tot = 0, ext = 0;
        unsigned long long fsum = 0;
        int lab = 0;
        for (int w = 0; w < kWaves; w++) {
            tot += lds.ep_count[w];
            ext += lds.ep_extra[w];
            fsum += lds.ep_fe[w];
            lab |= lds.ep_lab[w];
        }
        lds.ep_base = (tot | ext) ? add_agent(wq.out_len, tot + ext) : 0u;
        if (track && fsum) {
            __hip_atomic_fetch_add(&ctl->fedges[(step + 1) % 3], fsum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&ctl->visited_edges, fsum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lab) atomicMax(&ctl->max_depth, depth);
    }
    __syncthreads();
    unsigned off = lds.ep_base;
    for (int w = 0; w < wave; w++) off += lds.ep_count[w];
    const unsigned cnt = (unsigned) wq.count;
    const unsigned lim = off < wq.cap ? min(cnt, wq.cap - off) : 0u;
    if (lim < cnt && lane == 0) {
        st_agent(wq.overflow, 1u);
        st_agent(wq.overflow_step, 1u);
    }
    for (unsigned i = (unsigned) lane; i < lim; i += kWave) {
        wq.out[off + i] = wq.buf[i];
        if (PRED) wq.pout[off + i] = wq.pbuf[i];
    }
    wq.count = 0;
}

__device__ __forceinline__ unsigned long long ld64_agent(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Generic phases
// ---------------------------------------------------------------------------

// Tile phase of a gathering step.  row_of(i, beg, end, par) turns input
// element i (< n) into the adjacency row to gather (end == beg for none);
// edge(ok, u, par) consumes each neighbour, wave-converged.
template <bool PRED, typename RowOf, typename Edge>
__device__ __forceinline__ void tile_phase(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n,
                                           const HubList &hub, Lds<PRED> &lds, RowOf &&row_of, Edge &&edge)
{
    TileGrab tg = tile_grab(s, n, lds);
    while (true) {
        unsigned base = next_tile(&s.ctl->steal[c.step & 1], lds, tg) * kBlock;
        if (base >= n) break;
        unsigned i = base + threadIdx.x;
        int beg = 0, end = 0, par = -1;
        if (i < n) row_of(i, beg, end, par);
        gather_rows<PRED>(beg, end, par, g.col, s.wave_t, s.cta_t, hub, lds, edge);
    }
}

// Hub phase: every workgroup takes hub_chunk(n)-edge chunks of the rows that
// the tile phase published, until all chunks are done.  Must follow a grid
// barrier (or kernel boundary) after the tile phase.
template <bool PRED, typename Chunk>
__device__ __forceinline__ void hub_phase_chunks(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n,
                                                 Lds<PRED> &lds, bool marked, Chunk &&run_chunk)
{
    unsigned long long a = __hip_atomic_load(&s.ctl->hub_alloc[c.step & 1], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    unsigned entries = (unsigned) (a >> 32);
    if (entries > s.hub_cap) entries = s.hub_cap;
    if (entries == 0) return;
    const HubEntry *list = s.hubs[c.step & 1];
    HubEntry lastent = list[entries - 1];
    const int chunk = hub_chunk(n, marked);
    unsigned total = lastent.first_chunk + (unsigned) ((lastent.end - lastent.beg + chunk - 1) / chunk);
    // guided grabs of consecutive chunks, as for tiles: a level of many hub
    // chunks must not queue on the one chunk counter
    TileGrab tg{0u, 0u, max(1u, min((unsigned) s.grab_max, total / (gridDim.x * 8u))), 0u};
    unsigned hint = 0;  // thread 0: entry of this workgroup's previous chunk
    while (true) {
        unsigned ch = next_tile(&s.ctl->hub_steal[c.step & 1], lds, tg);
        if (ch >= total) break;
        if (threadIdx.x == 0) {
            const unsigned lo = hub_entry_of(list, entries, ch, hint);
            hint = lo;
            HubEntry e = list[lo];
            int b = e.beg + (int) (ch - e.first_chunk) * chunk;
            lds.cta_beg = b;
            lds.cta_end = min(e.end, b + chunk);
            lds.cta_par = e.par;
        }
        __syncthreads();
        run_chunk(lds.cta_beg, lds.cta_end, lds.cta_par);
    }
}


template <bool PRED, typename Edge>
__device__ __forceinline__ void hub_phase(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds,
                                          Edge &&edge)
{
    hub_phase_chunks<PRED>(g, s, c, n, lds, false, [&](int b, int e, int pv) {
        for (int k = b; k < e; k += kBlock) {
            int idx = k + threadIdx.x;
            bool ok = idx < e;
            int u = ok ? __builtin_nontemporal_load(g.col + idx) : -1;
            edge(ok, u, pv);
        }
    });
}

// ---------------------------------------------------------------------------
// Step bodies.  All threads of the workgroup execute them together.
// sub == 0 runs the tile phase, sub == 1 the hub phase.  The bodies and the
// level loop's other parts are force-inlined: one left out of line takes the
// kernel's argument structs by address, and the compiler then copies them to
// scratch and reads every field from there for the whole kernel.
// ---------------------------------------------------------------------------
template <bool PRED, bool EXACT>
__device__ __forceinline__ void body_expand_contract(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds,
                                     int sub, bool hubs, unsigned *next_bits = nullptr, bool from_bitmap = false)
{
    WaveQueue<PRED> wq = make_queue<PRED>(s, c.step, lds);
    const int *in = s.q[c.step & 1];
    const int next_depth = c.iteration + 1;
    bool labeled = false;
    unsigned fe = 0;
    auto edge = [&](bool ok, int u, int par) {
        bool fresh = ok && claim<EXACT>(s.visited, s.depth, u, s.coherent_filter);
        if (fresh) {
            s.depth[u] = next_depth;
            if (PRED) s.pred[u] = par;
            labeled = true;
            if (s.track_edges) fe += (unsigned) (g.row[u + 1] - g.row[u]);
            // a clear frontier bitmap becomes the next frontier's, exactly
            if (next_bits)
                __hip_atomic_fetch_or(next_bits + (u >> 5), 1u << (u & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        wq.push(fresh, u, par);
    };
    if (sub == 0 && from_bitmap) {
        // The frontier is frontier bitmap fbm[c.iteration & 1], as a
        // bottom-up level left it: each workgroup takes kBlock-word tiles
        // of it, clears them and the other bitmap's words (so both are
        // clear after the step, as a top-down level needs them), and
        // gathers one member per thread per round.
        const HubList hub = hub_list(s, c.step, hubs, n);
        unsigned *bm = s.fbm[c.iteration & 1];
        unsigned *other = s.fbm[(c.iteration + 1) & 1];
        const unsigned words = ((unsigned) g.nodes + 31) >> 5;
        for (unsigned w0 = blockIdx.x * kBlock; w0 < words; w0 += gridDim.x * kBlock) {
            const unsigned w = w0 + threadIdx.x;
            unsigned bits = 0u;
            if (w < words) {
                bits = bm[w];
                bm[w] = 0u;
                other[w] = 0u;
            }
            while (__syncthreads_or(bits != 0u)) {
                int beg = 0, end = 0, par = -1;
                if (bits) {
                    const int v = (int) ((w << 5) + (unsigned) __builtin_ctz(bits));
                    bits &= bits - 1u;
                    beg = g.row[v];
                    end = g.row[v + 1];
                    par = v;
                }
                gather_rows<PRED>(beg, end, par, g.col, s.wave_t, s.cta_t, hub, lds, edge);
            }
        }
    } else if (sub == 0) {
        tile_phase<PRED>(g, s, c, n, hub_list(s, c.step, hubs, n), lds, [&](unsigned i, int &beg, int &end, int &par) {
            int v = ld_queue(in + i);
            beg = g.row[v];
            end = g.row[v + 1];
            par = v;
        }, edge);
    } else {
        hub_phase<PRED>(g, s, c, n, lds, edge);
    }
    finish_step<PRED>(wq, lds, s.ctl, c.step, labeled, next_depth, fe, s.track_edges != 0);
}

// Bottom-up step of a direction-optimizing search (Beamer, Asanovic,
// Patterson, "Direction-Optimizing Breadth-First Search", SC'12): every
// unvisited vertex scans its row for a neighbour in the current frontier
// bitmap fbm[d & 1] and stops at the first one (in row order).
//
// The level is bound by each wave's chain of dependent loads (row bounds,
// neighbours, frontier words), not by bytes.  Each wave owns one contiguous
// run of 64-vertex groups (static partition: no tile counter, no workgroup
// barrier; scrambled ids balance the runs) and walks it in two stages:
//   1. the visited words and row bounds of kBuGroups groups at a time
//      (prefetched one step ahead) give the vertices still to search with a
//      non-empty row; they are queued in the wave's LDS -- on later levels
//      only a few lanes of a group have work;
//   2. every 128 queued vertices are scanned with all lanes busy, two per
//      lane: four neighbours of each per round, all eight loads of a round
//      in flight; rows of kBuLong+ neighbours get one such round and then a
//      wave-wide scan of the rest (a high-degree vertex usually finds a
//      frontier neighbour among its first four).
// Parents found are collected as bits in LDS per window of kBuWindow groups
// and written once per word: visited |= found, next frontier = found (the
// next-frontier word is overwritten in full, which also clears it).
constexpr int kBuLong = 64;
constexpr int kBuGroups = 2;
constexpr int kBuQueue = 192;   // queued vertices per wave (a two-entry scan and a group more)
constexpr int kBuWindow = 32;   // groups per window of found bits (2 words each)
static_assert(3 * kBuQueue + 2 * kBuWindow <= kWaveQueueCap, "bottom-up LDS use fits the wave's staging area");

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool PRED, bool EXACT>
__device__ __forceinline__ void body_bottom_up(const GraphArgs &g, const SearchArgs &s, Cursor c, Lds<PRED> &lds)
{
    WaveQueue<PRED> wq = make_queue<PRED>(s, c.step, lds);
    const unsigned *cur = s.fbm[c.iteration & 1];
    unsigned *nxt = s.fbm[(c.iteration + 1) & 1];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int next_depth = c.iteration + 1;
    const unsigned words = ((unsigned) g.nodes + 31) >> 5;
    bool labeled = false;
    unsigned fe = 0;
    int found = 0;
    int *qv = lds.wq[wave];
    int *qb = qv + kBuQueue;
    int *qe = qb + kBuQueue;
    unsigned *fbits = reinterpret_cast<unsigned *>(qe + kBuQueue);

    const unsigned groups = ((unsigned) g.nodes + kWave - 1) / kWave;
    const unsigned nwaves = gridDim.x * kWaves;
    const unsigned gw = blockIdx.x * kWaves + (unsigned) wave;
    const unsigned per = (groups + nwaves - 1) / nwaves;
    const unsigned g_begin = gw * per;
    const unsigned g_end = min(groups, (gw + 1) * per);

    // Scans the first cnt (1..128, wave-uniform) queued vertices of the
    // window whose first visited word is word0: lane l takes entries l and
    // l + 64, whose rounds of four neighbours run with all loads in flight.
    auto scan = [&](int cnt, unsigned word0) {
        int v[2], beg[2], end[2], pos[2], lim[2], parent[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = lane + h * kWave;
            v[h] = -1;
            beg[h] = 0;
            end[h] = 0;
            if (i < cnt) {
                v[h] = qv[i];
                beg[h] = qb[i];
                end[h] = qe[i];
            }
            pos[h] = beg[h];
            lim[h] = end[h] - beg[h] < kBuLong ? end[h] : min(end[h], beg[h] + 4);
            parent[h] = -1;
        }
        while ((parent[0] < 0 && pos[0] < lim[0]) || (parent[1] < 0 && pos[1] < lim[1])) {
            int u[8];
            unsigned fw[8];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    u[h * 4 + j] = parent[h] < 0 && pos[h] + j < lim[h] ? g.rcol[pos[h] + j] : -1;
#pragma unroll
            for (int k = 0; k < 8; k++) fw[k] = u[k] >= 0 ? cur[u[k] >> 5] : 0u;
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int j = 3; j >= 0; j--)  // the first hit in row order wins
                    if (u[h * 4 + j] >= 0 && ((fw[h * 4 + j] >> (u[h * 4 + j] & 31)) & 1u)) parent[h] = u[h * 4 + j];
            pos[0] += 4;
            pos[1] += 4;
        }
        // rows of kBuLong+ neighbours with no parent among their first four:
        // a wave-wide scan of the rest, one row at a time
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint64_t lm = ballot(parent[h] < 0 && lim[h] < end[h]);
            while (lm) {
                const int l = uniform(__builtin_ctzll(lm));
                lm &= lm - 1;
                const int b = readlane(lim[h], l), e = readlane(end[h], l);
                int hit = -1;
                for (int q = b; q < e; q += kWave) {
                    const int idx = q + lane;
                    const int uu = idx < e ? g.rcol[idx] : -1;
                    const uint64_t hm = ballot(uu >= 0 && ((cur[uu >> 5] >> (uu & 31)) & 1u));
                    if (hm) {
                        hit = readlane(uu, uniform(__builtin_ctzll(hm)));
                        break;
                    }
                }
                if (lane == l) parent[h] = hit;
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (parent[h] >= 0) {
                s.depth[v[h]] = next_depth;
                if (PRED) s.pred[v[h]] = parent[h];
                fe += (unsigned) (end[h] - beg[h]);
                labeled = true;
                found++;
                __hip_atomic_fetch_or(&fbits[((unsigned) v[h] >> 5) - word0], 1u << (v[h] & 31), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        }
    };
    // Moves queue entries [from, from + n) to the front.
    auto shift = [&](int from, int n) {
        int v = 0, b = 0, e = 0;
        if (lane < n) {
            v = qv[from + lane];
            b = qb[from + lane];
            e = qe[from + lane];
        }
        wave_lds_sync();
        if (lane < n) {
            qv[lane] = v;
            qb[lane] = b;
            qe[lane] = e;
        }
        wave_lds_sync();
    };

    unsigned vis_n[kBuGroups];
    int beg_n[kBuGroups], end_n[kBuGroups];
    auto load_groups = [&](unsigned g0, unsigned g_stop) {
#pragma unroll
        for (int k = 0; k < kBuGroups; k++) {
            const int vv = (int) ((g0 + (unsigned) k) * kWave + (unsigned) lane);
            const bool ok = g0 + (unsigned) k < g_stop && vv < g.nodes;
            vis_n[k] = ok ? s.visited[(unsigned) vv >> 5] : ~0u;
            beg_n[k] = ok ? g.rrow[vv] : 0;
            end_n[k] = ok ? g.rrow[vv + 1] : 0;
        }
    };
    if (g_begin < g_end) load_groups(g_begin, g_end);
    for (unsigned w_begin = g_begin; w_begin < g_end; w_begin += kBuWindow) {
        const unsigned w_end = min(g_end, w_begin + kBuWindow);
        const unsigned word0 = 2 * w_begin;
        fbits[lane] = 0u;  // 64 lanes, 2 * kBuWindow words
        wave_lds_sync();
        int qn = 0;
        for (unsigned gi = w_begin; gi < w_end; gi += kBuGroups) {
            unsigned vis[kBuGroups];
            int beg[kBuGroups], end[kBuGroups];
#pragma unroll
            for (int k = 0; k < kBuGroups; k++) {
                vis[k] = vis_n[k];
                beg[k] = beg_n[k];
                end[k] = end_n[k];
            }
            // the next groups of this run (the window boundary does not stop the pipeline)
            if (gi + kBuGroups < g_end) load_groups(gi + kBuGroups, g_end);
            bool todo[kBuGroups];
#pragma unroll
            for (int k = 0; k < kBuGroups; k++) {
                const int v = (int) ((gi + (unsigned) k) * kWave + (unsigned) lane);
                todo[k] = gi + (unsigned) k < w_end && v < g.nodes && ((vis[k] >> (v & 31)) & 1u) == 0 &&
                          end[k] > beg[k];
            }
            if (!EXACT) {
                // Best-effort top-down levels may have lost visited bits to
                // racing read-modify-writes; the label stays the authority.
                int dep[kBuGroups];
#pragma unroll
                for (int k = 0; k < kBuGroups; k++)
                    dep[k] = todo[k] ? s.depth[(gi + (unsigned) k) * kWave + (unsigned) lane] : -1;
#pragma unroll
                for (int k = 0; k < kBuGroups; k++) todo[k] = todo[k] && dep[k] == -1;
            }
#pragma unroll
            for (int k = 0; k < kBuGroups; k++) {
                const uint64_t m = ballot(todo[k]);
                if (todo[k]) {
                    const int p = qn + lanes_below(m);
                    qv[p] = (int) ((gi + (unsigned) k) * kWave + (unsigned) lane);
                    qb[p] = beg[k];
                    qe[p] = end[k];
                }
                qn += __popcll(m);
                if (qn >= 2 * kWave) {
                    wave_lds_sync();
                    scan(2 * kWave, word0);
                    shift(2 * kWave, qn - 2 * kWave);
                    qn -= 2 * kWave;
                }
            }
        }
        wave_lds_sync();
        while (qn > 0) {
            const int cnt = min(qn, 2 * kWave);
            scan(cnt, word0);
            if (qn > cnt) shift(cnt, qn - cnt);
            qn -= cnt;
        }
        wave_lds_sync();
        // the window's visited and next-frontier words (the wave owns them)
        const unsigned w = word0 + (unsigned) lane;
        if (lane < 2 * (int) (w_end - w_begin) && w < words) {
            const unsigned bits = fbits[lane];
            if (bits) s.visited[w] |= bits;
            nxt[w] = bits;
        }
        wave_lds_sync();
    }
    // The level's output is the bitmap nxt; only its size goes to the queue
    // length (queue_from_bitmap materialises the queue if the search turns
    // top-down again).
    finish_step<PRED>(wq, lds, s.ctl, c.step, labeled, next_depth, fe, true, found);
}

// Writes the members of frontier bitmap fbm[c.iteration & 1] into the input
// queue of step c (a bottom-up level records its output only as a bitmap),
// clearing both frontier bitmaps on the way.
// Grid-wide; a grid barrier or kernel boundary must follow before the queue
// is read.  Reservations come from ctl->compact, which the previous step
// zeroed.
template <bool PRED>
__device__ __forceinline__ void queue_from_bitmap(const GraphArgs &g, const SearchArgs &s, Cursor c, Lds<PRED> &lds)
{
    unsigned *bm = s.fbm[c.iteration & 1];
    unsigned *other = s.fbm[(c.iteration + 1) & 1];
    int *out = s.q[c.step & 1];
    const unsigned words = ((unsigned) g.nodes + 31) >> 5;
    const unsigned stride = gridDim.x * kBlock;
    for (unsigned w0 = blockIdx.x * kBlock; w0 < words; w0 += stride) {
        const unsigned w = w0 + threadIdx.x;
        const unsigned bits = w < words ? bm[w] : 0u;
        // top-down levels leave both bitmaps clear (see frontier_bitmap_pass)
        if (w < words) {
            bm[w] = 0u;
            other[w] = 0u;
        }
        int total = 0;
        int excl = block_exclusive_sum(__popc(bits), lds.ep_scan, &total);
        if (threadIdx.x == 0) lds.ep_base = total ? add_agent(&s.ctl->compact, (unsigned) total) : 0u;
        __syncthreads();
        unsigned pos = lds.ep_base + (unsigned) excl;
        for (unsigned b = bits; b; b &= b - 1) {
            if (pos < s.cap) out[pos] = (int) ((w << 5) + (unsigned) __builtin_ctz(b));
            pos++;
        }
        __syncthreads();
    }
}

template <bool PRED, bool EXACT>
__device__ __forceinline__ void body_contract_expand(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds,
                                     int sub, bool hubs)
{
    WaveQueue<PRED> wq = make_queue<PRED>(s, c.step, lds);
    const int *in = s.q[c.step & 1];
    const int *pin = s.pq[c.step & 1];
    bool labeled = false;
    auto edge = [&](bool ok, int w, int par) { wq.push(ok, w, par); };
    if (sub == 0) {
        tile_phase<PRED>(g, s, c, n, hub_list(s, c.step, hubs, n), lds, [&](unsigned i, int &beg, int &end, int &par) {
            int u = ld_queue(in + i);
            int p = PRED ? ld_queue(pin + i) : -1;
            if (history_pass<PRED>(lds, u) && claim<EXACT>(s.visited, s.depth, u, s.coherent_filter)) {
                s.depth[u] = c.iteration;
                if (PRED) s.pred[u] = p;
                beg = g.row[u];
                end = g.row[u + 1];
                par = u;
                labeled = true;
            }
        }, edge);
    } else {
        hub_phase<PRED>(g, s, c, n, lds, edge);
    }
    finish_step<PRED>(wq, lds, s.ctl, c.step, labeled, c.iteration, 0u, false);
}

template <bool PRED, bool EXACT>
__device__ __forceinline__ void body_expand(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds, int sub,
                            bool hubs)
{
    WaveQueue<PRED> wq = make_queue<PRED>(s, c.step, lds);
    const int *in = s.q[c.step & 1];
    auto edge = [&](bool ok, int w, int par) { wq.push(ok, w, par); };
    if (sub == 0) {
        tile_phase<PRED>(g, s, c, n, hub_list(s, c.step, hubs, n), lds, [&](unsigned i, int &beg, int &end, int &par) {
            int v = ld_queue(in + i);
            beg = g.row[v];
            end = g.row[v + 1];
            par = v;
        }, edge);
    } else {
        hub_phase<PRED>(g, s, c, n, lds, edge);
    }
    finish_step<PRED>(wq, lds, s.ctl, c.step, false, 0, 0u, false);
}

// ---------------------------------------------------------------------------
// Chained compaction (config compaction = 2) of the steps whose output per
// tile is bounded by the tile (contract: claim, filter: bitmask cull).  The
// unit of the chain is a grab (the consecutive tiles one tile-counter atomic
// hands a workgroup).  Pass 1 decides every item of the grab and keeps one
// ballot mask per wave and tile in LDS; wave 0 sums the masks, publishes
// the grab's count in its 64-bit status word (epoch, state, value) and finds
// the grab's offset in the output queue by a decoupled look-back over the
// earlier grabs' words (prim::lookback_prefix) -- a single-pass grid scan in
// place of the reference's per-tile atomicAdd reservation
// (util/scan/cooperative_scan.cuh:309-343) and its multi-kernel scan passes.
// Pass 2 re-reads the kept items (still in cache) and writes them at the
// offset plus their rank, so the output keeps the input order.  Grabs are
// taken in increasing order and publish their count before looking back,
// so a grab waits only for grabs already taken, each at most one pass 1
// away.  The last grab stores the output length.
// ---------------------------------------------------------------------------
constexpr unsigned kChainGrabMax = kWave;  // tiles per grab: one lane each in the count scan
static_assert(kChainGrabMax * kWaves * 2 <= kWaves * kWaveQueueCap, "grab masks fit the staging area");

template <bool PRED, bool EXACT, bool FILTER>
__device__ __forceinline__ void chained_step(const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds,
                                             bool &labeled)
{
    const int *in = s.q[c.step & 1];
    const int *pin = s.pq[c.step & 1];
    int *out = s.q[(c.step + 1) & 1];
    int *pout = PRED ? s.pq[(c.step + 1) & 1] : nullptr;
    // the wave-queue staging is idle: ballot masks [tile][wave]
    unsigned long long *masks = reinterpret_cast<unsigned long long *>(&lds.wq[0][0]);
    int *tile_off = &lds.sc_excl[0][0];
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const unsigned tiles = (n + kBlock - 1) / kBlock;
    // (grabs in increasing order, as the look-back needs: always dynamic)
    const unsigned g = min(tile_grab_size(s, tiles), kChainGrabMax);
    const unsigned grabs = (tiles + g - 1) / g;
    const unsigned epoch = (s.scan_epoch + (unsigned) c.step) & 0x3FFFFFFFu;
    while (true) {
        const unsigned first = next_item(&s.ctl->steal[c.step & 1], lds, g);
        if (first >= tiles) break;
        const unsigned cnt = min(g, tiles - first);
        for (unsigned k = 0; k < cnt; k++) {
            const unsigned i = (first + k) * kBlock + threadIdx.x;
            bool keep = false;
            if (i < n) {
                const int u = ld_queue(in + i);
                if (FILTER) {
                    if (u >= 0) keep = ((s.visited[u >> 5] >> (u & 31)) & 1u) == 0;
                } else {
                    const int p = PRED ? ld_queue(pin + i) : -1;
                    keep = u >= 0 && history_pass<PRED>(lds, u) && claim<EXACT>(s.visited, s.depth, u, s.coherent_filter);
                    if (keep) {
                        s.depth[u] = c.iteration;
                        if (PRED) s.pred[u] = p;
                        labeled = true;
                    }
                }
            }
            const unsigned long long m = ballot(keep);
            if (lane == 0) masks[k * kWaves + wave] = m;
        }
        __syncthreads();
        if (wave == 0) {
            int cnt_l = 0;
            if ((unsigned) lane < cnt)
                for (int w = 0; w < kWaves; w++) cnt_l += __popcll(masks[lane * kWaves + w]);
            const int incl = wave_inclusive_sum(cnt_l);
            if ((unsigned) lane < cnt) tile_off[lane] = incl - cnt_l;
            const unsigned agg = (unsigned) readlane(incl, kWave - 1);
            const unsigned j = first / g;
            unsigned prefix = 0;
            if (j == 0) {
                if (lane == 0) prim::publish(s.tile_status, 0, epoch, prim::kStatePrefix, agg);
            } else {
                if (lane == 0) prim::publish(s.tile_status, j, epoch, prim::kStateAggregate, agg);
                prefix = prim::lookback_prefix(s.tile_status, j, epoch);
                if (lane == 0) prim::publish(s.tile_status, j, epoch, prim::kStatePrefix, prefix + agg);
            }
            if (lane == 0) {
                lds.ep_base = prefix;
                if (j == grabs - 1) st_agent(&s.ctl->qlen[(c.step + 1) % 3], prefix + agg);
            }
        }
        __syncthreads();
        const unsigned prefix = lds.ep_base;
        bool lost = false;
        for (unsigned k = 0; k < cnt; k++) {
            const unsigned long long m = masks[k * kWaves + wave];
            if (!((m >> lane) & 1ull)) continue;
            unsigned rank = (unsigned) tile_off[k] + (unsigned) __popcll(m & ((1ull << lane) - 1ull));
            for (int w = 0; w < wave; w++) rank += (unsigned) __popcll(masks[k * kWaves + w]);
            const unsigned i = (first + k) * kBlock + threadIdx.x;
            const unsigned at = prefix + rank;
            if (at < s.cap) {
                out[at] = ld_queue(in + i);
                if (PRED) pout[at] = FILTER ? ld_queue(pin + i) : -1;
            } else {
                lost = true;
            }
        }
        if (lost) {
            st_agent(&s.ctl->overflow, 1u);
            st_agent(&s.ctl->ovf[(c.step + 1) % 3], 1u);
        }
        // (next_item's barrier orders these mask reads before the next grab's writes)
    }
}

template <bool PRED, bool EXACT>
__device__ __forceinline__ void body_contract(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds)
{
    (void) g;
    WaveQueue<PRED> wq = make_queue<PRED>(s, c.step, lds);
    const int *in = s.q[c.step & 1];
    const int *pin = s.pq[c.step & 1];
    bool labeled = false;
    if (s.tile_status) {
        chained_step<PRED, EXACT, false>(s, c, n, lds, labeled);
    } else {
        TileGrab tg = tile_grab(s, n, lds);
        while (true) {
            unsigned base = next_tile(&s.ctl->steal[c.step & 1], lds, tg) * kBlock;
            if (base >= n) break;
            unsigned i = base + threadIdx.x;
            int u = -1, p = -1;
            if (i < n) {
                u = ld_queue(in + i);
                if (PRED) p = ld_queue(pin + i);
            }
            bool keep = u >= 0 && history_pass<PRED>(lds, u) && claim<EXACT>(s.visited, s.depth, u, s.coherent_filter);
            if (keep) {
                s.depth[u] = c.iteration;
                if (PRED) s.pred[u] = p;
                labeled = true;
            }
            wq.push(keep, u, -1);
        }
    }
    finish_step<PRED>(wq, lds, s.ctl, c.step, labeled, c.iteration, 0u, false);
}

template <bool PRED, bool EXACT>
__device__ __forceinline__ void body_filter(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds)
{
    (void) g;
    WaveQueue<PRED> wq = make_queue<PRED>(s, c.step, lds);
    const int *in = s.q[c.step & 1];
    const int *pin = s.pq[c.step & 1];
    if (s.tile_status) {
        bool unused = false;
        chained_step<PRED, EXACT, true>(s, c, n, lds, unused);
    } else {
        TileGrab tg = tile_grab(s, n, lds);
        while (true) {
            unsigned base = next_tile(&s.ctl->steal[c.step & 1], lds, tg) * kBlock;
            if (base >= n) break;
            unsigned i = base + threadIdx.x;
            int u = -1, p = -1;
            if (i < n) {
                u = ld_queue(in + i);
                if (PRED) p = ld_queue(pin + i);
            }
            // Bitmask only: the history hash is reserved for claims, because a
            // hit there must mean the vertex has already been claimed or found
            // visited by this workgroup.
            bool keep = false;
            if (u >= 0) {
                unsigned w = s.visited[u >> 5];
                keep = ((w >> (u & 31)) & 1u) == 0;
            }
            wq.push(keep, u, p);
        }
    }
    finish_step<PRED>(wq, lds, s.ctl, c.step, false, 0, 0u, false);
}

__device__ __forceinline__ bool gathers(int phase)
{
    return phase == kPhaseExpandContract || phase == kPhaseContractExpand || phase == kPhaseExpand;
}

template <bool PRED, bool EXACT>
__device__ __forceinline__ void run_body(int phase, const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n,
                                         Lds<PRED> &lds, int sub, bool hubs)
{
    switch (phase) {
    case kPhaseExpandContract: body_expand_contract<PRED, EXACT>(g, s, c, n, lds, sub, hubs); break;
    case kPhaseContractExpand: body_contract_expand<PRED, EXACT>(g, s, c, n, lds, sub, hubs); break;
    case kPhaseContract: body_contract<PRED, EXACT>(g, s, c, n, lds); break;
    case kPhaseExpand: body_expand<PRED, EXACT>(g, s, c, n, lds, sub, hubs); break;
    case kPhaseBottomUp: body_bottom_up<PRED, EXACT>(g, s, c, lds); break;
    default: body_filter<PRED, EXACT>(g, s, c, n, lds); break;
    }
}

// The phases a persistent loop can run (the others are not compiled into
// its kernel, which keeps their registers out of its budget).
template <bool PRED, bool EXACT, int LOOP>
__device__ __forceinline__ void loop_body(int phase, const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n,
                                          Lds<PRED> &lds, int sub, bool hubs, unsigned *next_bits = nullptr,
                                          bool from_bitmap = false)
{
    if constexpr (LOOP == kLoopEC) {
        body_expand_contract<PRED, EXACT>(g, s, c, n, lds, sub, hubs);
    } else if constexpr (LOOP == kLoopCE) {
        // (expand: the step that ends a contract-expand fast path)
        if (phase == kPhaseExpand) body_expand<PRED, EXACT>(g, s, c, n, lds, sub, hubs);
        else body_contract_expand<PRED, EXACT>(g, s, c, n, lds, sub, hubs);
    } else if constexpr (LOOP == kLoopDO) {
        if (phase == kPhaseBottomUp) body_bottom_up<PRED, EXACT>(g, s, c, lds);
        else body_expand_contract<PRED, EXACT>(g, s, c, n, lds, sub, hubs, next_bits, from_bitmap);
    } else {
        if (phase == kPhaseContract) body_contract<PRED, EXACT>(g, s, c, n, lds);
        else if (phase == kPhaseExpand) body_expand<PRED, EXACT>(g, s, c, n, lds, sub, hubs);
        else body_filter<PRED, EXACT>(g, s, c, n, lds);
    }
}

// Two grid-wide passes that turn the vertex queue of step c into the
// frontier bitmap fbm[c.iteration & 1] (zero, then set); a grid barrier or
// kernel boundary must separate them.  A persistent search skips the first:
// its prologue clears both bitmaps and queue_from_bitmap clears them again
// whenever the search turns top-down, so they are clear at every top-down
// level.
__device__ __forceinline__ void frontier_bitmap_pass(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n,
                                                     int pass)
{
    unsigned *bm = s.fbm[c.iteration & 1];
    const unsigned stride = gridDim.x * kBlock;
    if (pass == 0) {
        const unsigned words = ((unsigned) g.nodes + 31) >> 5;
        for (unsigned w = blockIdx.x * kBlock + threadIdx.x; w < words; w += stride) bm[w] = 0u;
    } else {
        const int *in = s.q[c.step & 1];
        for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
            int u = ld_queue(in + i);
            __hip_atomic_fetch_or(&bm[u >> 5], 1u << (u & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// Mark-and-compact form of a large top-down step (expand-contract and the
// top-down levels of direction-optimizing search).  The claims of a big
// level are bound by atomic throughput (about 27 G/s on this part, whatever
// the scope), while plain random byte stores run at about 84 G/s.  Such a
// level therefore only marks its discoveries: every edge to a vertex whose
// visited bit is clear (the bits do not change during the pass) stores 1
// into marks[u] -- duplicates merge, as every writer stores the same byte
// -- and, for predecessors, the edge's source into pred[u] (any writer
// leaves a valid parent).  After a grid barrier one streaming pass turns
// the marks of each 32-vertex word into visited bits, labels and queue
// entries (in vertex order) and clears them.  The pass and its barrier cost
// about 10-15 us, so small levels keep the atomic claims.
// ---------------------------------------------------------------------------
constexpr unsigned long long kMarkEdges = 1ull << 20;  // fewest frontier edges for the marked form

// fe counts the frontier's edges with multiplicity (best-effort levels may
// queue a vertex more than once).
__device__ __forceinline__ bool marked_step(const SearchArgs &s, unsigned long long fe)
{
    return s.marks != nullptr && fe >= kMarkEdges && fe >= (unsigned long long) s.nodes / 2;
}

template <bool PRED, bool EXACT>
__device__ __forceinline__ void body_expand_mark(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, Lds<PRED> &lds,
                                 int sub, bool hubs)
{
    // An edge marks its target when the target's visited bit is clear and,
    // in best-effort mode, its label is still -1: that mode may have lost
    // visited bits of labelled vertices, whose predecessors a mark would
    // overwrite.
    const int *in = s.q[c.step & 1];
    if (sub == 0) {
        // Tiles of kBlock frontier vertices: rows of the grid tier are
        // published, the rest of the tile's edges are spread evenly over the
        // workgroup through an LDS degree scan (as in the LDS fast path),
        // four consecutive edges per thread per round with every load of a
        // round in flight -- instead of the gather tiers' one chain per
        // round of 64 edges.
        const HubList hub = hub_list(s, c.step, hubs, n, true);
        int *tbeg = lds.wq[0];  // the wave-queue staging is idle in a marked step
        int *tpar = lds.wq[1];
        int *toff = lds.wq[2];
        static_assert(kWaveQueueCap >= kBlock && kWaves >= 3, "tile tables fit the staging area");
        TileGrab tg = tile_grab(s, n, lds);
        while (true) {
            const unsigned base = next_tile(&s.ctl->steal[c.step & 1], lds, tg) * kBlock;
            if (base >= n) break;
            const unsigned i = base + threadIdx.x;
            int v = -1, beg = 0, end = 0;
            if (i < n) {
                v = ld_queue(in + i);
                beg = g.row[v];
                end = g.row[v + 1];
            }
            if (publish_hub_row(beg, end, v, hub)) end = beg;
            int total = 0;
            const int off = block_exclusive_sum(end - beg, lds.ep_scan, &total);
            // (the scan's barriers order the previous tile's reads before these writes)
            tbeg[threadIdx.x] = beg;
            tpar[threadIdx.x] = v;
            toff[threadIdx.x] = off;
            __syncthreads();
            for (int b0 = 0; b0 < total; b0 += kBlock * 4) {
                const int e0 = b0 + (int) threadIdx.x * 4;
                int r = 0;
                if (e0 < total) {  // last row with off <= e0 (it has edges)
                    int lo = 0, hi = kBlock - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (toff[mid] <= e0) lo = mid;
                        else hi = mid - 1;
                    }
                    r = lo;
                }
                int at[4], pv[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int e = e0 + j;
                    at[j] = -1;
                    pv[j] = -1;
                    if (e < total) {
                        while (r + 1 < kBlock && toff[r + 1] <= e) r++;
                        at[j] = tbeg[r] + (e - toff[r]);
                        pv[j] = tpar[r];
                    }
                }
                int u[4];
                unsigned w[4];
#pragma unroll
                for (int j = 0; j < 4; j++) u[j] = at[j] >= 0 ? g.col[at[j]] : -1;
#pragma unroll
                for (int j = 0; j < 4; j++) w[j] = u[j] >= 0 ? s.visited[u[j] >> 5] : ~0u;
                bool todo[4];
#pragma unroll
                for (int j = 0; j < 4; j++) todo[j] = u[j] >= 0 && !((w[j] >> (u[j] & 31)) & 1u);
                if (!EXACT) {
                    int dep[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) dep[j] = todo[j] ? s.depth[u[j]] : -1;
#pragma unroll
                    for (int j = 0; j < 4; j++) todo[j] = todo[j] && dep[j] == -1;
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (todo[j]) {
                        s.marks[u[j]] = 1;
                        if (PRED) s.pred[u[j]] = pv[j];
                    }
                }
            }
        }
        return;
    }
    // Grid-tier chunks, four workgroup-widths of edges per round with every
    // load of a round in flight together (the per-edge chain -- column,
    // visited word, mark -- is the cost of these levels, not bytes).
    constexpr int U = 4;
    hub_phase_chunks<PRED>(g, s, c, n, lds, true, [&](int b, int e, int pv) {
        for (int k = b; k < e; k += U * kBlock) {
            int u[U];
            unsigned w[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                const int idx = k + j * kBlock + (int) threadIdx.x;
                u[j] = idx < e ? g.col[idx] : -1;
            }
#pragma unroll
            for (int j = 0; j < U; j++) w[j] = u[j] >= 0 ? s.visited[u[j] >> 5] : ~0u;
            bool todo[U];
#pragma unroll
            for (int j = 0; j < U; j++) todo[j] = u[j] >= 0 && !((w[j] >> (u[j] & 31)) & 1u);
            if (!EXACT) {
                int dep[U];
#pragma unroll
                for (int j = 0; j < U; j++) dep[j] = todo[j] ? s.depth[u[j]] : -1;
#pragma unroll
                for (int j = 0; j < U; j++) todo[j] = todo[j] && dep[j] == -1;
            }
#pragma unroll
            for (int j = 0; j < U; j++) {
                if (todo[j]) {
                    s.marks[u[j]] = 1;
                    if (PRED) s.pred[u[j]] = pv;
                }
            }
        }
    });
}

// Bits of the four 0/1 bytes of x.
__device__ __forceinline__ unsigned byte_bits(unsigned x)
{
    return (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
}

// Grid-wide; follows the marking pass behind a grid barrier, and a grid
// barrier must follow it.  Appends to the queue of step c + 1 like the
// claiming form (one reservation per workgroup and word batch), in vertex
// order.  With fbm_out set, also writes every word of that frontier bitmap
// (the next step's frontier, exactly: a direction-optimizing search that
// turns bottom-up next needs no bitmap pass).
template <bool PRED, bool EXACT>
__device__ __forceinline__ void mark_compact(const GraphArgs &g, const SearchArgs &s, Cursor c, Lds<PRED> &lds, unsigned *fbm_out)
{
    const unsigned words = ((unsigned) g.nodes + 31) >> 5;
    const int next_depth = c.iteration + 1;
    int *out = s.q[(c.step + 1) & 1];
    // the new vertices of a word batch, staged in the (idle) wave-queue LDS
    int *stage = &lds.wq[0][0];
    constexpr int kStage = kWaves * kWaveQueueCap;
    unsigned fe = 0;
    bool labeled = false;
    bool over = false;
    for (unsigned w0 = blockIdx.x * kBlock; w0 < words; w0 += gridDim.x * kBlock) {
        const unsigned w = w0 + threadIdx.x;
        unsigned bits = 0;
        if (w < words) {
            uint4 *m = reinterpret_cast<uint4 *>(s.marks + 32ull * w);
            const uint4 lo = m[0], hi = m[1];
            bits = byte_bits(lo.x) | byte_bits(lo.y) << 4 | byte_bits(lo.z) << 8 | byte_bits(lo.w) << 12 |
                   byte_bits(hi.x) << 16 | byte_bits(hi.y) << 20 | byte_bits(hi.z) << 24 | byte_bits(hi.w) << 28;
            if (bits) {
                m[0] = make_uint4(0u, 0u, 0u, 0u);
                m[1] = make_uint4(0u, 0u, 0u, 0u);
                const unsigned vis = s.visited[w];
                bits &= ~vis;
                if (!EXACT) {
                    // best-effort levels may have lost bits of labelled vertices
                    for (unsigned b = bits; b; b &= b - 1) {
                        const int v = (int) ((w << 5) + (unsigned) __builtin_ctz(b));
                        if (s.depth[v] != -1) bits &= ~(1u << (v & 31));
                    }
                }
                if (bits) s.visited[w] = vis | bits;
            }
            if (fbm_out) fbm_out[w] = bits;
        }
        int total = 0;
        const int excl = block_exclusive_sum(__popc(bits), lds.ep_scan, &total);
        if (total == 0) continue;  // block-uniform
        if (threadIdx.x == 0) lds.ep_base = add_agent(&s.ctl->qlen[(c.step + 1) % 3], (unsigned) total);
        labeled = labeled || bits != 0;
        if (total <= kStage) {
            // labels, queue entries and degrees spread over the workgroup,
            // four vertices per thread per round with their row loads in
            // flight together (a thread's own word may hold 32)
            int k = excl;
            for (unsigned b = bits; b; b &= b - 1) stage[k++] = (int) ((w << 5) + (unsigned) __builtin_ctz(b));
            __syncthreads();
            const unsigned base = lds.ep_base;
            for (int i0 = 4 * (int) threadIdx.x; i0 < total; i0 += 4 * kBlock) {
                int v[4], r0[4], r1[4];
#pragma unroll
                for (int j = 0; j < 4; j++) v[j] = i0 + j < total ? stage[i0 + j] : -1;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    r0[j] = v[j] >= 0 && s.track_edges ? g.row[v[j]] : 0;
                    r1[j] = v[j] >= 0 && s.track_edges ? g.row[v[j] + 1] : 0;
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (v[j] < 0) continue;
                    s.depth[v[j]] = next_depth;
                    const unsigned pos = base + (unsigned) (i0 + j);
                    if (pos < s.cap) out[pos] = v[j];
                    else over = true;
                    fe += (unsigned) (r1[j] - r0[j]);
                }
            }
        } else {
            __syncthreads();
            unsigned pos = lds.ep_base + (unsigned) excl;
            for (unsigned b = bits; b; b &= b - 1) {
                const int v = (int) ((w << 5) + (unsigned) __builtin_ctz(b));
                s.depth[v] = next_depth;
                if (s.track_edges) fe += (unsigned) (g.row[v + 1] - g.row[v]);
                if (pos < s.cap) out[pos] = v;
                else over = true;
                pos++;
            }
        }
        __syncthreads();  // ep_base and the stage are reused by the next batch
    }
    if (over) {
        st_agent(&s.ctl->overflow, 1u);
        st_agent(&s.ctl->ovf[(c.step + 1) % 3], 1u);
    }
    // frontier-edge, visited-edge and depth counters: one update per workgroup
    WaveQueue<PRED> wq = make_queue<PRED>(s, c.step, lds);
    finish_step<PRED>(wq, lds, s.ctl, c.step, labeled, next_depth, fe, s.track_edges != 0);
}

// Direction choice of a direction-optimizing search at the top of a level
// (grid-uniform: every input is read after the level barrier).
//
// ve is the caller's running sum of fedges over the steps up to this one.
// The global visited_edges total must not be read here: workgroups already
// in this level's body keep adding to it, and a workgroup reading it late
// could choose the other direction -- and then a different sequence of grid
// barriers than the rest of the grid.
// Direction rule: mf = the frontier's edges, n = its vertices, ve = edges
// visited so far, prev_mf = the previous frontier's edges.  A top-down
// search turns bottom-up only while its frontier grows (Beamer et al.'s
// growing phase): near the end of a high-diameter search the unexplored
// edges fall below alpha times a shrinking frontier's, and a bottom-up
// level there would turn straight back.
__device__ __forceinline__ int direction_rule(const SearchArgs &s, int phase, unsigned long long mf, unsigned n,
                                              unsigned long long ve, unsigned long long prev_mf)
{
    unsigned long long mu = s.total_edges > ve ? s.total_edges - ve : 0ull;
    if (phase == kPhaseExpandContract) {
        if (mf > prev_mf && mf * (unsigned long long) s.do_alpha > mu) return kPhaseBottomUp;
    } else if ((unsigned long long) n * (unsigned long long) s.do_beta < (unsigned long long) s.nodes) {
        return kPhaseExpandContract;
    }
    return phase;
}


// Whether a level is small enough for the single-workgroup fast path.
// fe: the step's frontier edges (used when they are tracked).
__device__ __forceinline__ bool fast_level(const SearchArgs &s, Cursor c, unsigned n, unsigned long long fe)
{
    if (s.fast_t < 0 || n > (unsigned) s.fast_t || c.phase == kPhaseBottomUp) return false;
    if (s.track_edges && fe > 32ull * (unsigned) s.fast_t) return false;
    return true;
}

__device__ __forceinline__ bool fast_level(const SearchArgs &s, Cursor c, unsigned n)
{
    return fast_level(s, c, n, s.track_edges ? ld64_agent(&s.ctl->fedges[c.step % 3]) : 0ull);
}

__device__ __forceinline__ Cursor advance(int loop, Cursor c)
{
    c.step++;
    if (loop != kLoopTP) {
        // (a contract-expand loop returns to contract-expand after the
        // expand step that ends its fast path: the expanded members were
        // labelled with c.iteration, their neighbours get c.iteration + 1)
        c.iteration++;
        if (loop == kLoopCE) c.phase = kPhaseContractExpand;
    } else {
        if (c.phase == kPhaseContract) c.phase = kPhaseExpand;
        else if (c.phase == kPhaseExpand) c.phase = kPhaseFilter;
        else { c.phase = kPhaseContract; c.iteration++; }
    }
    return c;
}

// Counter maintenance for step c (one thread): zero the counters the next
// step appends to / steals from, and account the input length.
__device__ __forceinline__ void prepare_step(Control *ctl, Cursor c, unsigned n)
{
    st_agent(&ctl->qlen[(c.step + 2) % 3], 0u);
    st_agent(&ctl->steal[(c.step + 1) & 1], 0u);
    st_agent(&ctl->hub_steal[(c.step + 1) & 1], 0u);
    st_agent(&ctl->compact, 0u);
    st_agent(&ctl->ovf[(c.step + 2) % 3], 0u);
    __hip_atomic_store(&ctl->hub_alloc[(c.step + 1) & 1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctl->fedges[(c.step + 2) % 3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ctl->total_queued += n;
    if (c.step < kLevelLog) {
        ctl->level_len[c.step] = n;
        ctl->level_clock[c.step] = wall_clock64();
    }
}

__device__ __forceinline__ unsigned frontier_length(const SearchArgs &s, int step)
{
    unsigned n = ld_agent(&s.ctl->qlen[step % 3]);
    return n < s.cap ? n : s.cap;
}

__device__ __forceinline__ bool hubs_pending(const SearchArgs &s, int step)
{
    return __hip_atomic_load(&s.ctl->hub_alloc[step & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull;
}

// ---------------------------------------------------------------------------
// LDS-resident fast path of the vertex-frontier loops (expand-contract and
// the top-down levels of direction-optimizing search).  Block 0 alone runs
// consecutive small levels while the rest of the grid waits at a grid
// barrier.  The frontier, its members' row starts and degrees stay in LDS
// between levels, and the global queue and counters are written once, when
// the path ends.  The level's edges are spread evenly over the threads, each
// taking kFastK consecutive edges through the exclusive degree scan, and each
// memory trip of a batch is issued for all kFastK edges before the first
// result is used: column index, visited word, claim, row bounds -- four
// dependent trips per batch, and a level of up to kBlock * kFastK edges is
// one batch.  A high-diameter graph spends its whole search here.
// ---------------------------------------------------------------------------
constexpr int kFastCap = 1024;  // frontier members kept in LDS
constexpr int kFastK = 4;       // consecutive edges per thread per batch
// Most edges of a level run by the path: a batch of kBlock * kFastK edges
// costs about one quarter of a grid-wide step, so a few batches still win.
constexpr unsigned long long kFastEdges = 4ull * kBlock * kFastK;
static_assert(kFastCap <= 4 * kBlock, "the degree scan takes four members per thread");

// Direct levels: when every member's degree is small, groups of lanes take
// whole rows (a thread up to kFastDirect edges) with no degree scan and no
// search for each edge's member.
constexpr int kFastDirect = 8;
static_assert(kFastDirect == 8, "fast_direct_lg spans one to eight lanes per member");

// log2 of the lanes per member of a direct level whose members' largest
// degree is maxd (30: too large for one)
__device__ __forceinline__ int fast_direct_lg(int maxd)
{
    return maxd <= 1 ? 0 : maxd <= 2 ? 1 : maxd <= 4 ? 2 : maxd <= kFastDirect ? 3 : 30;
}

__device__ __forceinline__ bool fast_direct(int maxd, unsigned n)
{
    const int lg = fast_direct_lg(maxd);
    return lg < 30 && ((int) n << lg) <= kBlock * kFastDirect;
}

struct FastFront {
    int v[2][kFastCap];
    int beg[2][kFastCap];
    int deg[2][kFastCap];        // degree on append; the level's exclusive scan in place
    int maxdeg[2];               // largest degree of each buffer's members
    unsigned count;              // next-frontier appends
    unsigned dups;               // two-phase: further edges to vertices claimed in the level
    unsigned long long fe;       // next frontier's edges
    int labeled;
    int scan[kWaves];            // per-wave sums of the degree scan
    unsigned long long wfe[kWaves];
    int wlab[kWaves];
    int wmax[kWaves];
};

// FASTV: the fast path's copy of the visited bitmap takes the rest of the
// workgroup's LDS (gfx950: 160 KiB) once the frontier buffers and 1 KiB for
// the level loop's other variables are placed: graphs of up to
// kFastVisitedWords * 32 vertices.  Larger graphs launch the instantiation
// without it, which keeps the smaller LDS footprint.
constexpr int kLdsBytes = 163840;
constexpr int kFastVisitedWords = ((kLdsBytes - 1024 - (int) sizeof(FastFront)) / 4) & ~63;
static_assert(kFastVisitedWords * 32 == BFS_FAST_VISITED_MAX, "the header states the fast path's LDS bitmap size");

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <bool FASTV>
struct FastLds {
    FastFront f;
    __attribute__((aligned(16))) unsigned vbits[FASTV ? kFastVisitedWords : 4];
};

// The level loop's LDS: the step bodies' staging and the fast path's
// frontier are never live at once (the fast path starts after a grid
// barrier and re-initialises the step state when it ends).
template <bool PRED, bool FASTV>
union LoopLds {
    Lds<PRED> step;
    FastLds<FASTV> fast;
};

// Exclusive scan of F.deg[b][0, n) in place; returns the sum to every thread.
__device__ int fast_scan(FastFront &F, int b, unsigned n)
{
    const unsigned t0 = 4u * threadIdx.x;
    int d[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        d[k] = t0 + k < n ? F.deg[b][t0 + k] : 0;
        sum += d[k];
    }
    int total;
    int off = block_exclusive_sum(sum, F.scan, &total);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (t0 + k < n) F.deg[b][t0 + k] = off;
        off += d[k];
    }
    __syncthreads();
    return total;
}

// Claims of one thread's edge slots: at[k] = column position of slot k (or
// -1), pv[k] = its row's owner.  Claims the unvisited neighbours, labels
// them, and returns the mask of the slots k that claimed u[k] (row
// [ub[k], ud[k])).  Each memory trip (column entry, visited word, claim,
// row bounds) is issued for every slot before the first result is used.
//
// LOCAL: the calling workgroup is the only one touching the labels and the
// visited words (block 0's fast path, between two grid barriers).  Its
// claims are then workgroup-scope atomics, done in this XCD's L2, where
// agent-scope ones must be coherent across the eight XCDs; the barrier
// before the path invalidated stale lines, and the one after it writes the
// path's lines back.  Column entries are read through the caches: the next
// level's rows sit next to this level's on a lattice.  With vbits (LOCAL
// only) the claims are LDS atomics on the path's copy of the visited
// bitmap: no pre-read and no memory trip between a column entry and the
// claimed vertex's row bounds.
// fast_claim_vertices: the same for vertices u[k] (-1: none) already at hand.
template <bool PRED, bool EXACT, bool LOCAL, int S>
__device__ __forceinline__ unsigned fast_claim_vertices(const GraphArgs &g, const SearchArgs &s, int next_depth,
                                                        const int (&pv)[S], const int (&u)[S], int (&ub)[S],
                                                        int (&ud)[S], bool ldsv, unsigned *vbits)
{
    unsigned old[S];
    if (LOCAL && ldsv) {
#pragma unroll
        for (int k = 0; k < S; k++) {
            old[k] = ~0u;
            if (u[k] >= 0)
                old[k] = __hip_atomic_fetch_or(vbits + (u[k] >> 5), 1u << (u[k] & 31), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    } else {
        // visited words (a set bit is final), then claims of the clear ones
        unsigned cur[S];
#pragma unroll
        for (int k = 0; k < S; k++) {
            cur[k] = ~0u;
            if (u[k] >= 0) {
                unsigned *w = s.visited + (u[k] >> 5);
                cur[k] = s.coherent_filter ? __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *w;
            }
        }
        // The claims' results are read only after all are issued (a use
        // next to its atomic would wait out each round trip in turn).
#pragma unroll
        for (int k = 0; k < S; k++) {
            old[k] = ~0u;
            const unsigned bit = 1u << (u[k] & 31);
            if (u[k] >= 0 && !(cur[k] & bit)) {
                // agent scope unless LOCAL: other workgroups claim
                // concurrently, and another XCD's L2 may hold this word from
                // before an earlier step's claims
                if (LOCAL)
                    old[k] = __hip_atomic_fetch_or(s.visited + (u[k] >> 5), bit, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                    old[k] = __hip_atomic_fetch_or(s.visited + (u[k] >> 5), bit, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    // (the empty asm keeps the compiler from folding a use into the
    // atomic's own block)
#pragma unroll
    for (int k = 0; k < S; k++) asm volatile("" : "+v"(old[k]));
    unsigned fresh = 0;
#pragma unroll
    for (int k = 0; k < S; k++)
        if (!((old[k] >> (u[k] & 31)) & 1u)) fresh |= 1u << k;
    if (!EXACT) {
        // best-effort culling elsewhere may have lost bits of labelled vertices
        int dep[S];
#pragma unroll
        for (int k = 0; k < S; k++) dep[k] = ((fresh >> k) & 1) ? s.depth[u[k]] : -1;
#pragma unroll
        for (int k = 0; k < S; k++)
            if (dep[k] != -1) fresh &= ~(1u << k);
    }
#pragma unroll
    for (int k = 0; k < S; k++) {
        ub[k] = 0;
        ud[k] = 0;
        if ((fresh >> k) & 1) {
            s.depth[u[k]] = next_depth;
            if (PRED) s.pred[u[k]] = pv[k];
            ub[k] = g.row[u[k]];
            ud[k] = g.row[u[k] + 1];
        }
    }
    return fresh;
}

template <bool PRED, bool EXACT, bool LOCAL, int S>
__device__ __forceinline__ unsigned fast_claim_slots(const GraphArgs &g, const SearchArgs &s, int next_depth,
                                                     const int (&at)[S], const int (&pv)[S], int (&u)[S],
                                                     int (&ub)[S], int (&ud)[S], bool ldsv = false,
                                                     unsigned *vbits = nullptr)
{
#pragma unroll
    for (int k = 0; k < S; k++)
        u[k] = at[k] < 0 ? -1 : LOCAL ? g.col[at[k]] : __builtin_nontemporal_load(g.col + at[k]);
    return fast_claim_vertices<PRED, EXACT, LOCAL, S>(g, s, next_depth, pv, u, ub, ud, ldsv, vbits);
}

// kFastK consecutive edges from e0 (one thread's share of a batch) of an
// LDS-resident frontier: offs = the exclusive scan of its n members' degrees
// (total edges in all), begs = their row starts, vs = the members.  Returns
// fast_claim_slots' mask; pv[k] = the owner of slot k's row.
template <bool PRED, bool EXACT, bool LOCAL>
__device__ __forceinline__ unsigned fast_edges(const GraphArgs &g, const SearchArgs &s, int next_depth,
                                               const int *offs, const int *begs, const int *vs, unsigned n,
                                               int total, int e0, int (&u)[kFastK], int (&pv)[kFastK],
                                               int (&ub)[kFastK], int (&ud)[kFastK], bool ldsv = false,
                                               unsigned *vbits = nullptr)
{
    // member holding edge e0: the last i with offs[i] <= e0 (it has edges)
    int i = 0;
    if (e0 < total) {
        int lo = 0, hi = (int) n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= e0) lo = mid;
            else hi = mid - 1;
        }
        i = lo;
    }
    // column positions first (LDS only): a loop between two global
    // loads makes the compiler wait for the first
    int at[kFastK];
#pragma unroll
    for (int k = 0; k < kFastK; k++) {
        const int e = e0 + k;
        at[k] = -1;
        pv[k] = -1;
        if (e < total) {
            while (i + 1 < (int) n && offs[i + 1] <= e) i++;
            at[k] = begs[i] + (e - offs[i]);
            pv[k] = vs[i];
        }
    }
    return fast_claim_slots<PRED, EXACT, LOCAL, kFastK>(g, s, next_depth, at, pv, u, ub, ud, ldsv, vbits);
}

// Appends one thread's claimed slots to LDS buffer dst (one LDS reservation
// per wave); members past kFastCap go to the global queue `spill` at the
// same positions.  Wave-converged.
template <int S>
__device__ __forceinline__ void fast_append(const SearchArgs &s, FastFront &F, int dst, int *spill, unsigned fresh,
                                            const int (&u)[S], const int (&ub)[S], const int (&ud)[S],
                                            unsigned long long &fe, bool &labeled, int &dmax)
{
    const int cnt = __popc(fresh);
    const int incl = wave_inclusive_sum(cnt);
    const int wtot = readlane(incl, kWave - 1);
    if (!wtot) return;
    unsigned base = 0;
    if (lane_id() == 0) base = atomicAdd(&F.count, (unsigned) wtot);
    unsigned p = (unsigned) readlane((int) base, 0) + (unsigned) (incl - cnt);
#pragma unroll
    for (int k = 0; k < S; k++) {
        if ((fresh >> k) & 1) {
            const int d = ud[k] - ub[k];
            fe += (unsigned) d;
            dmax = max(dmax, d);
            labeled = true;
            if (p < (unsigned) kFastCap) {
                F.v[dst][p] = u[k];
                F.beg[dst][p] = ub[k];
                F.deg[dst][p] = d;
            } else if (p < s.cap) {
                spill[p] = u[k];
            }
            p++;
        }
    }
}

// Folds the threads' edge sums, label flags and largest degrees of the
// members just appended to buffer dst into F.fe, F.labeled, F.maxdeg[dst].
// Every thread calls it.
__device__ __forceinline__ void fast_fold(FastFront &F, int dst, unsigned long long fe, bool labeled, int dmax)
{
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const int lo = wave_inclusive_sum((int) (fe & 0xFFFFu));
    const int hi = wave_inclusive_sum((int) (fe >> 16));
    const int wmax = wave_inclusive_max_nonneg(dmax);
    const bool any = ballot(labeled) != 0;
    if (lane == kWave - 1) {
        F.wfe[wave] = ((unsigned long long) (unsigned) hi << 16) + (unsigned long long) (unsigned) lo;
        F.wlab[wave] = any ? 1 : 0;
        F.wmax[wave] = wmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long f = 0;
        int l = 0, dm = 0;
        for (int w = 0; w < kWaves; w++) {
            f += F.wfe[w];
            l |= F.wlab[w];
            dm = max(dm, F.wmax[w]);
        }
        F.fe = f;
        F.labeled = l;
        F.maxdeg[dst] = dm;
    }
    __syncthreads();
}

// One level from LDS buffer src (n members, 0 < n <= kFastCap) into buffer
// src ^ 1; members past kFastCap go to the global queue `spill` (the one the
// step after the path reads) at the same positions.  When the members' rows fit kFastDirect slots per thread
// (every lattice level), each member's row goes to a group of consecutive
// lanes, one edge per lane; otherwise the level's edges are spread evenly
// over the threads through the degree scan, kFastK per thread per batch.
// Every thread calls it.
// count_dups (direct levels only): also counts in F.dups the edges, other
// than the claiming ones, whose targets the level claimed.
template <bool PRED, bool EXACT>
__device__ __forceinline__ void fast_lds_level(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, int src,
                                               FastFront &F, bool ldsv, unsigned *vbits, int *spill,
                                               bool count_dups = false)
{
    const int dst = src ^ 1;
    const int next_depth = c.iteration + 1;
    const int maxd = F.maxdeg[src];
    if (threadIdx.x == 0) {
        F.count = 0;
        F.dups = 0;
    }
    unsigned long long fe = 0;
    bool labeled = false;
    int dmax = 0;
    // Direct level: 2^lg consecutive lanes per member (2^lg >= its largest
    // degree), lane k of a group taking edge k, so that the column entries
    // one load instruction reads are whole rows side by side; kFastDirect
    // such passes cover up to kBlock * kFastDirect / 2^lg members.
    const int lg = fast_direct_lg(maxd);
    if (fast_direct(maxd, n)) {
        __syncthreads();  // the resets above precede every append
        int at[kFastDirect], pv[kFastDirect];
        const int k = (int) threadIdx.x & ((1 << lg) - 1);
        const int i0 = (int) threadIdx.x >> lg;
#pragma unroll
        for (int j = 0; j < kFastDirect; j++) {
            const int i = i0 + ((j * kBlock) >> lg);
            at[j] = -1;
            pv[j] = -1;
            if (i < (int) n && k < F.deg[src][i]) {
                at[j] = F.beg[src][i] + k;
                pv[j] = F.v[src][i];
            }
        }
        int u[kFastDirect], ub[kFastDirect], ud[kFastDirect];
        const unsigned fresh =
            fast_claim_slots<PRED, EXACT, true, kFastDirect>(g, s, next_depth, at, pv, u, ub, ud, ldsv, vbits);
        fast_append<kFastDirect>(s, F, dst, spill, fresh, u, ub, ud, fe, labeled, dmax);
        if (count_dups) {
            fast_fold(F, dst, fe, labeled, dmax);
            // (after the fold's barriers every label of the level is
            // visible to the workgroup)
            int dep[kFastDirect];
#pragma unroll
            for (int j = 0; j < kFastDirect; j++)
                dep[j] = u[j] >= 0 && !((fresh >> j) & 1) ? s.depth[u[j]] : -1;
            int d = 0;
#pragma unroll
            for (int j = 0; j < kFastDirect; j++) d += dep[j] == next_depth ? 1 : 0;
            const int wsum = wave_inclusive_sum(d);
            if (lane_id() == kWave - 1 && wsum) atomicAdd(&F.dups, (unsigned) wsum);
            __syncthreads();
            return;
        }
    } else {
        const int total = fast_scan(F, src, n);  // (its barriers order the resets before the appends)
        for (int b0 = 0; b0 < total; b0 += kBlock * kFastK) {
            int u[kFastK], pv[kFastK], ub[kFastK], ud[kFastK];
            const unsigned fresh =
                fast_edges<PRED, EXACT, true>(g, s, next_depth, F.deg[src], F.beg[src], F.v[src], n, total,
                                              b0 + (int) threadIdx.x * kFastK, u, pv, ub, ud, ldsv, vbits);
            fast_append<kFastK>(s, F, dst, spill, fresh, u, ub, ud, fe, labeled, dmax);
        }
    }
    fast_fold(F, dst, fe, labeled, dmax);
}

// Spread step: a top-down level of few members (n <= kFastCap) with more
// edges than the fast path takes -- typically the first level after the
// source, a handful of hubs.  The tile phase would hand each row to one
// workgroup, or publish the long ones for a hub phase behind another grid
// barrier; here every workgroup loads the whole frontier into LDS, scans its
// degrees and walks an equal share of the level's edges (kBlock * kFastK at
// a time, strided over the grid), so the level is one phase and one grid
// barrier whatever its row lengths.  Each batch reserves its queue slots
// with one atomic per workgroup.  next_bits as for body_expand_contract.
// Every thread of every workgroup calls it; F must not be live.
template <bool PRED, bool EXACT>
__device__ __forceinline__ void spread_step(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, FastFront &F,
                            unsigned *next_bits)
{
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const int *in = s.q[c.step & 1];
    for (unsigned i = threadIdx.x; i < n; i += kBlock) {
        const int v = ld_queue(in + i);
        const int b = g.row[v];
        F.v[0][i] = v;
        F.beg[0][i] = b;
        F.deg[0][i] = g.row[v + 1] - b;
    }
    __syncthreads();
    const int total = fast_scan(F, 0, n);
    const int next_depth = c.iteration + 1;
    Control *ctl = s.ctl;
    int *out = s.q[(c.step + 1) & 1];
    int *pout = PRED ? s.pq[(c.step + 1) & 1] : nullptr;
    unsigned long long fe = 0;
    bool labeled = false;
    const int stride = (int) gridDim.x * kBlock * kFastK;
    for (int b0 = (int) blockIdx.x * kBlock * kFastK; b0 < total; b0 += stride) {
        int u[kFastK], pv[kFastK], ub[kFastK], ud[kFastK];
        const unsigned fresh = fast_edges<PRED, EXACT, false>(g, s, next_depth, F.deg[0], F.beg[0], F.v[0], n, total,
                                                       b0 + (int) threadIdx.x * kFastK, u, pv, ub, ud);
        int got;
        const int off = block_exclusive_sum(__popc(fresh), F.scan, &got);
        // (every thread read the previous batch's base before the scan's
        // first barrier)
        if (threadIdx.x == 0 && got) F.count = add_agent(&ctl->qlen[(c.step + 1) % 3], (unsigned) got);
        __syncthreads();
        const unsigned p0 = F.count + (unsigned) off;
        bool lost = false;
#pragma unroll
        for (int k = 0; k < kFastK; k++) {
            const unsigned p = p0 + (unsigned) __popc(fresh & ((1u << k) - 1u));
            if ((fresh >> k) & 1) {
                fe += (unsigned) (ud[k] - ub[k]);
                if (p < s.cap) {
                    out[p] = u[k];
                    if (PRED) pout[p] = pv[k];
                } else {
                    lost = true;
                }
            }
        }
        labeled |= fresh != 0;
        if (lost) {
            st_agent(&ctl->overflow, 1u);
            st_agent(&ctl->ovf[(c.step + 1) % 3], 1u);
        }
        if (next_bits) {
#pragma unroll
            for (int k = 0; k < kFastK; k++)
                if ((fresh >> k) & 1)
                    __hip_atomic_fetch_or(next_bits + (u[k] >> 5), 1u << (u[k] & 31), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // one update per workgroup of the next frontier's edges and the depth
    const int lo = wave_inclusive_sum((int) (fe & 0xFFFFu));
    const int hi = wave_inclusive_sum((int) (fe >> 16));
    const bool any = ballot(labeled) != 0;
    if (lane == kWave - 1) {
        F.wfe[wave] = ((unsigned long long) (unsigned) hi << 16) + (unsigned long long) (unsigned) lo;
        F.wlab[wave] = any ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long f = 0;
        int l = 0;
        for (int w = 0; w < kWaves; w++) {
            f += F.wfe[w];
            l |= F.wlab[w];
        }
        if (s.track_edges && f) {
            __hip_atomic_fetch_add(&ctl->fedges[(c.step + 1) % 3], f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&ctl->visited_edges, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (l) atomicMax(&ctl->max_depth, next_depth);
    }
}

// Runs LDS fast levels from step c (input: n <= kFastCap members in the
// global queue) until a level's output is empty, too large for the path, or
// (direction-optimizing) calls for bottom-up.  Leaves the global queue,
// counters and the fast_* hand-off exactly as the ordinary steps would have.
// `fv`: visited-edge sum through step c, fe_in: step c's frontier edges
// (see direction_rule).  Block 0 only; every thread calls it.
template <bool PRED, bool EXACT, bool FASTV>
__device__ __forceinline__ void fast_lds_path(const GraphArgs &g, const SearchArgs &s, Cursor c, unsigned n, int loop,
                              unsigned long long fv, unsigned long long fe_in, FastLds<FASTV> &L, bool wait_others,
                              const BarrierShape &bshape)
{
    FastFront &F = L.f;
    const int *in = s.q[c.step & 1];
    // FASTV: the path's claims go to an LDS copy of the visited bitmap (no
    // other workgroup reads or writes visited words until the grid barrier
    // that ends the path); it is written back before that barrier.
    const bool ldsv = FASTV && s.nodes <= kFastVisitedWords * 32;
    const int vwords = (s.nodes + 31) >> 5;
    if (ldsv) {
        // 16-byte chunks, kLoad per thread in flight together
        constexpr int kLoad = 16;
        const u32x4 *from = reinterpret_cast<const u32x4 *>(s.visited);
        u32x4 *to = reinterpret_cast<u32x4 *>(L.vbits);
        const int chunks = vwords >> 2;
        for (int base = 0; base < chunks; base += kBlock * kLoad) {
            u32x4 t[kLoad];
#pragma unroll
            for (int j = 0; j < kLoad; j++) {
                const int i = base + j * kBlock + (int) threadIdx.x;
                if (i < chunks) t[j] = from[i];
            }
#pragma unroll
            for (int j = 0; j < kLoad; j++) {
                const int i = base + j * kBlock + (int) threadIdx.x;
                if (i < chunks) to[i] = t[j];
            }
        }
        for (int i = 4 * chunks + (int) threadIdx.x; i < vwords; i += kBlock) L.vbits[i] = s.visited[i];
    }
    int src = 0;
    unsigned long long ve_add = 0, queued = 0;
    int maxd = -1;
    unsigned n2 = 0;
    unsigned long long fe2 = 0;
    // frontier edges of the members and of the level before them
    unsigned long long cur_fe = fe_in, prev_fe = 0;
    const bool ce = loop == kLoopCE, tp = loop == kLoopTP;
    if (ce || tp) {
        // Contract-expand and two-phase enter at a step that contracts a
        // list of candidates (the previous level's neighbours or their
        // filtered remainder, with their parents).  Its claims, labelled
        // with the step's iteration as the step labels them, become the LDS
        // members; each level below then does the next contraction.
        // Contract-expand counts a step's input (the frontier edges of the
        // members before it) in total_queued; two-phase counts an expand
        // (the members), a filter (their edges) and a contraction (the
        // edges to vertices the level claims).
        if (threadIdx.x == 0) {
            F.count = 0;
            if (c.step < kLevelLog) {
                s.ctl->level_len[c.step] = n;
                s.ctl->level_clock[c.step] = wall_clock64();
            }
        }
        __syncthreads();
        queued += n;
        const int *pin = (c.step & 1) ? s.pq[1] : s.pq[0];
        static_assert(kFastCap <= 2 * kBlock, "two input items per thread");
        int u[2], pv[2], ub[2], ud[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const unsigned i = threadIdx.x + (unsigned) k * kBlock;
            u[k] = i < n ? ld_queue(in + i) : -1;
            pv[k] = PRED && i < n ? ld_queue(pin + i) : -1;
        }
        const unsigned fresh =
            fast_claim_vertices<PRED, EXACT, true, 2>(g, s, c.iteration, pv, u, ub, ud, ldsv, L.vbits);
        unsigned long long fe = 0;
        bool labeled = false;
        int dmax = 0;
        fast_append<2>(s, F, 0, (c.step & 1) ? s.q[0] : s.q[1], fresh, u, ub, ud, fe, labeled, dmax);
        fast_fold(F, 0, fe, labeled, dmax);
        n = F.count;  // (at most the input's n <= kFastCap: nothing spilled)
        fe2 = F.fe;
        if (F.labeled) maxd = c.iteration;
        if (n == 0) c = advance(loop, c);  // the next step's input is empty
    } else {
        int dmax = 0;
        for (unsigned i = threadIdx.x; i < n; i += kBlock) {
            const int v = ld_queue(in + i);
            const int b = g.row[v];
            const int d = g.row[v + 1] - b;
            F.v[0][i] = v;
            F.beg[0][i] = b;
            F.deg[0][i] = d;
            dmax = max(dmax, d);
        }
        dmax = wave_inclusive_max_nonneg(dmax);
        if (lane_id() == kWave - 1) F.wmax[threadIdx.x / kWave] = dmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            int dm = 0;
            for (int w = 0; w < kWaves; w++) dm = max(dm, F.wmax[w]);
            F.maxdeg[0] = dm;
        }
        __syncthreads();
    }
    while (n > 0) {
        // Two-phase counts a contraction's input -- every edge the filter
        // passes, i.e. every edge to a vertex the level claims -- which
        // direct levels count; the grid takes the others.
        if (tp && !fast_direct(F.maxdeg[src], n)) {
            n2 = n;
            break;
        }
        // (a contract-expand level is the contraction of step c + 1)
        const int logged = c.step + (ce ? 1 : 0);
        const unsigned long long step_input = ce ? fe2 : tp ? n + fe2 : n;
        if (threadIdx.x == 0 && logged < kLevelLog) {
            s.ctl->level_len[logged] = (unsigned) step_input;
            s.ctl->level_clock[logged] = wall_clock64();
        }
        queued += step_input;
        // the queue the step after the path reads if this level ends it
        // (members past kFastCap go there): that of step c + 1, or of the
        // expand step four steps on from a two-phase contraction
        int *spill = tp ? s.q[c.step & 1] : s.q[(c.step + 1) & 1];
        fast_lds_level<PRED, EXACT>(g, s, c, n, src, F, ldsv, L.vbits, spill, tp);
        n2 = F.count;
        fe2 = F.fe;
        if (F.labeled) maxd = c.iteration + 1;
        ve_add += fe2;
        src ^= 1;
        c = advance(loop, c);
        if (tp) {
            // expand and filter, then the contraction of the filter's output
            queued += (unsigned long long) n2 + F.dups;
            c = advance(loop, advance(loop, c));
        }
        bool go = n2 > 0 && n2 <= (unsigned) kFastCap && n2 <= (unsigned) s.fast_t && fe2 <= kFastEdges;
        if (go && loop == kLoopDO) {
            const unsigned long long fv2 = fv + fe2;
            if (direction_rule(s, c.phase, fe2, n2, fv2, cur_fe) != c.phase) go = false;
            else fv = fv2;
        }
        prev_fe = cur_fe;
        cur_fe = fe2;
        // a hybrid search leaves contract-expand when a step's input
        // exceeds the saturation threshold: the grid takes that step
        if (go && ce && s.sat_t >= 0 && fe2 > (unsigned long long) s.sat_t) go = false;
        n = n2;
        if (!go) break;
        __syncthreads();  // every thread has read F's results before the next level resets them
    }
    // Contract-expand hands its members to the grid as the input of an
    // expand step (cursor c, phase expand: their neighbours become step
    // c + 1's edge frontier), whose input the loop top adds to total_queued
    // although it is not a contract-expand step's.
    if (ce && n2 > 0) {
        c.phase = kPhaseExpand;
        queued -= min(n2, s.cap);
    }
    // Two-phase hands its members to the expand step after their
    // contraction, which counts them as its input.
    if (tp && n2 > 0) c = advance(loop, c);
    if (ldsv) {
        // (the last level's claims precede its closing workgroup barrier)
        const u32x4 *from = reinterpret_cast<const u32x4 *>(L.vbits);
        u32x4 *to = reinterpret_cast<u32x4 *>(s.visited);
        const int chunks = vwords >> 2;
        for (int i = threadIdx.x; i < chunks; i += kBlock) to[i] = from[i];
        for (int i = 4 * chunks + (int) threadIdx.x; i < vwords; i += kBlock) s.visited[i] = L.vbits[i];
    }
    // The hand-off below rewrites counters the other workgroups may still
    // be reading at the top of the step the path started from, unless they
    // have all arrived at the barrier that ends the path.
    if (wait_others) {
        if (threadIdx.x == 0) wait_other_arrivals(s.ctl, bshape);
        __syncthreads();
    }
    // The next step (c) reads its input from the global queue: copy the
    // LDS-resident members (those past kFastCap were written there already).
    const unsigned held = min(n2, (unsigned) kFastCap);
    int *q = s.q[c.step & 1];
    for (unsigned i = threadIdx.x; i < held && i < s.cap; i += kBlock) q[i] = F.v[src][i];
    if (threadIdx.x == 0) {
        Control *ctl = s.ctl;
        // what the previous step's prepare_step would have left for step c
        st_agent(&ctl->qlen[(c.step + 1) % 3], 0u);
        st_agent(&ctl->steal[c.step & 1], 0u);
        st_agent(&ctl->hub_steal[c.step & 1], 0u);
        st_agent(&ctl->compact, 0u);
        st_agent(&ctl->ovf[(c.step + 1) % 3], 0u);
        __hip_atomic_store(&ctl->hub_alloc[c.step & 1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->fedges[(c.step + 1) % 3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // step c's input
        st_agent(&ctl->qlen[c.step % 3], min(n2, s.cap));
        st_agent(&ctl->ovf[c.step % 3], n2 > s.cap ? 1u : 0u);
        if (n2 > s.cap) st_agent(&ctl->overflow, 1u);
        __hip_atomic_store(&ctl->fedges[c.step % 3], fe2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&ctl->visited_edges, ve_add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (maxd >= 0) atomicMax(&ctl->max_depth, maxd);
        ctl->total_queued += queued;
        st_agent(&ctl->fast_step, c.step);
        st_agent(&ctl->fast_iteration, c.iteration);
        st_agent(&ctl->fast_phase, c.phase);
        // excludes the step the loop top runs next (it adds it)
        __hip_atomic_store(&ctl->fast_vsum, fv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->fast_prev_fe, prev_fe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Search start and end inside a persistent launch (one launch per search).
// ---------------------------------------------------------------------------

// Every workgroup rewrites its share of the labels (with init_reset); block 0
// clears the status head and the level log and queues the source.  Only the
// labels and block 0's own words are written, so the first step may begin
// before a grid barrier as long as it reads nothing of the control block
// and claims nothing before one (the loop top takes step 0's values from
// search_start_values).
__device__ __forceinline__ void search_prologue(const GraphArgs &g, const SearchArgs &s, bool clear_bitmaps)
{
    const int src = s.init_src;
    if (clear_bitmaps) {
        // both frontier bitmaps, which are contiguous
        const int words = 2 * ((s.nodes + 31) >> 5);
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < words; i += gridDim.x * kBlock) s.fbm[0][i] = 0u;
    }
    if (s.init_reset) {
        const int words = (s.nodes + 31) >> 5;
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < s.nodes; i += gridDim.x * kBlock) {
            const bool is_src = s.init_label && i == src;
            s.depth[i] = is_src ? 0 : -1;
            if (s.pred) s.pred[i] = is_src ? -2 : -1;
            if (i < words) s.visited[i] = (s.init_label && (src >> 5) == i) ? 1u << (src & 31) : 0u;
        }
    }
    if (blockIdx.x != 0) return;
    Control *ctl = s.ctl;
    unsigned *head = reinterpret_cast<unsigned *>(ctl);
    for (int i = threadIdx.x; i < (int) (kControlHead / 4); i += kBlock) head[i] = 0u;
    for (int i = threadIdx.x; i < kLevelLog; i += kBlock) {
        ctl->level_len[i] = 0u;
        ctl->level_clock[i] = 0ull;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->qlen[0] = 1;
        ctl->max_depth = -1;
        s.q[0][0] = src;
        if (s.pq[0]) s.pq[0][0] = -2;
        if (s.init_label) {
            if (!s.init_reset) {
                s.depth[src] = 0;
                if (s.pred) s.pred[src] = -2;
                s.visited[src >> 5] |= 1u << (src & 31);
            }
            ctl->max_depth = 0;
            const unsigned long long deg = (unsigned long long) (g.row[src + 1] - g.row[src]);
            ctl->fedges[0] = deg;
            ctl->visited_edges = deg;
        }
    }
}

// Block 0 copies the status head to s.host_head when the level loop ends,
// so the host reads it after the launch without a copy of its own.  The
// loop ends at a loop top, after the grid barrier that completed the last
// step's counter updates, and the exit fields are block 0's own: the head
// is final (the instrumented cycle sums, added later, are not copied).
// Called by thread 0 of block 0.
__device__ __forceinline__ void publish_status(const SearchArgs &s)
{
    const unsigned *head = reinterpret_cast<const unsigned *>(s.ctl);
    unsigned *out = reinterpret_cast<unsigned *>(s.host_head);
    unsigned w[kControlHead / 4];
#pragma unroll
    for (int i = 0; i < (int) (kControlHead / 4); i++) w[i] = __hip_atomic_load(head + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int i = 0; i < (int) (kControlHead / 4); i++) out[i] = w[i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // system scope: the host reads it next
}

// ---------------------------------------------------------------------------
// Persistent level loop (cooperative launch).
// ---------------------------------------------------------------------------
// One instantiation per strategy (LOOP), so profiles attribute time to the
// strategy and the level loop carries no dead strategy branches.
// Most work of a step whose grid barrier leaves the L2 write-back to each
// XCD's last arriver, and the barrier of a launch from which XCD leaves
// are used.
constexpr unsigned long long kLightWork = 1ull << 19;
constexpr unsigned kXcdBarrierFrom = 32;

template <bool PRED, bool EXACT, int LOOP, bool FASTV>
__global__ __launch_bounds__(kBlock) void persistent_kernel(GraphArgs g, SearchArgs s, Cursor c)
{
    constexpr int loop = LOOP;
    __shared__ LoopLds<PRED, FASTV> shared;
    Lds<PRED> &lds = shared.step;
    init_lds(lds);
    unsigned long long t_begin = s.instrument ? wall_clock64() : 0, busy = 0;
    int reason = kExitDone;
    // Visited-edge sum up to the current step, kept by every workgroup from
    // counters that no workgroup changes during the step (fedges[step % 3])
    // -- see direction_rule.  fedges[0] holds the source's degree.
    unsigned long long vsum = 0;
    // the previous step's frontier edges (direction rule; grid-uniform too)
    unsigned long long prev_fe = 0;

    // With a folded search start, step 0's loop-top values come from the
    // graph (one source, its degree), not from the control block that block
    // 0 is still writing, and the step's first grid barrier orders the
    // label rewrite before any claim.
    bool start = s.init_src >= 0;
    // Grid barriers: leaves by block, and from the kXcdBarrierFrom-th
    // barrier of a long launch on (a high-diameter search) the XCDs, counted
    // from the records every workgroup left before its first barrier
    // (thread 0's shape is the one used; the count costs a few
    // microseconds, which a short launch does not win back).  After a light
    // phase (little written) only each XCD's last arriver then writes its L2
    // back; after a heavy one every workgroup still writes back its XCD's L2
    // before arriving, which spreads the write-back over the arrivals
    // instead of leaving it all to the last one.
    BarrierShape bshape = barrier_shape_by_block();
    bool shape_pending = gridDim.x >= 8u && gridDim.x <= (unsigned) kMaxXcdGrid;
    unsigned barriers = 0;
    if (shape_pending && threadIdx.x == 0) xcd_register(s.ctl);
    auto bar = [&](bool light) {
        BarrierShape b = bshape;
        b.shared = bshape.shared && light ? 1u : 0u;
        grid_barrier(s.ctl, b);
        if (shape_pending && ++barriers == kXcdBarrierFrom) {
            shape_pending = false;
            if (threadIdx.x == 0) bshape = xcd_shape(s.ctl);
        }
    };
    if (start) search_prologue(g, s, LOOP == kLoopDO);
    // Frontier bitmaps of a direction-optimizing loop (grid-uniform state):
    // which are not known to be clear, and the step whose frontier one of
    // them holds exactly (written by a marked step's compaction).
    unsigned fbm_dirty = start ? 0u : 3u;
    int fbm_exact_step = -1;

    // Every decision at the top of a step reads only values final before
    // the step began and unchanged during it (the input length, the
    // previous step's overflow flag, this step's frontier edges), so all
    // workgroups take the same branch and meet the same grid barriers.
    while (true) {
        unsigned n;
        unsigned long long fe = 0;
        bool bu_exit = false;  // bottom-up until this step
        if (start) {
            n = 1;
            if (s.track_edges && s.init_label) fe = (unsigned long long) (g.row[s.init_src + 1] - g.row[s.init_src]);
        } else {
            n = frontier_length(s, c.step);
            if (ld_agent(&s.ctl->ovf[c.step % 3])) { reason = kExitOverflow; break; }
            if (n == 0) { reason = kExitDone; break; }
            if (s.track_edges) fe = ld64_agent(&s.ctl->fedges[c.step % 3]);
        }
        // (an expand step that ends a contract-expand fast path runs first)
        if (loop == kLoopCE && c.phase != kPhaseExpand && s.sat_t >= 0 && n > (unsigned) s.sat_t) {
            reason = kExitSaturated;
            break;
        }
        if (loop == kLoopDO) {
            vsum += fe;
            int dir = direction_rule(s, c.phase, fe, n, vsum, prev_fe);
            prev_fe = fe;
            if (start && dir != c.phase) {
                bar(false);
                start = false;
            }
            if (dir == kPhaseBottomUp && c.phase != kPhaseBottomUp) {
                if (fbm_exact_step != c.step) {
                    if (fbm_dirty & (1u << (c.iteration & 1))) {
                        frontier_bitmap_pass(g, s, c, n, 0);
                        bar(false);
                    }
                    frontier_bitmap_pass(g, s, c, n, 1);
                    bar(false);
                }
                fbm_dirty = 3u;  // bottom-up levels write both
            } else if (dir != kPhaseBottomUp && c.phase == kPhaseBottomUp) {
                bu_exit = true;
            }
            c.phase = dir;
        }

        // Contract-expand and two-phase take the LDS path from a step that
        // contracts at most kFastCap items (never from the expand step the
        // path hands over to the grid with).
        const bool contract_lds = ((LOOP == kLoopCE && c.phase == kPhaseContractExpand) ||
                                   (LOOP == kLoopTP && c.phase == kPhaseContract)) &&
                                  n <= (unsigned) kFastCap;
        const bool fast = fast_level(s, c, n, fe) && gridDim.x > 1 && !(LOOP == kLoopCE && c.phase == kPhaseExpand) &&
                          ((LOOP != kLoopEC && LOOP != kLoopDO) || (n <= (unsigned) kFastCap && fe <= kFastEdges));
        // large top-down levels mark instead of claiming (grid-uniform: fe
        // is this step's frontier-edge count); few-member ones spread their
        // edges over the grid
        const bool mark =
            (LOOP == kLoopEC || LOOP == kLoopDO) && c.phase == kPhaseExpandContract && marked_step(s, fe);
        const bool spread = (LOOP == kLoopEC || LOOP == kLoopDO) && c.phase == kPhaseExpandContract && !mark &&
                            n <= (unsigned) kFastCap && s.spread;
        // The first top-down step after bottom-up levels reads its frontier
        // from the bitmap they left, unless its form needs the queue (the
        // fast path, spread and marked steps index it).
        const bool from_bitmap = bu_exit && !fast && !mark && !spread;
        if (bu_exit && !from_bitmap) {
            queue_from_bitmap<PRED>(g, s, c, lds);  // clears both bitmaps
            bar(false);
            fbm_dirty = 0u;
        }
        if (start && !fast) bar(false);
        const bool first = start;
        start = false;
        if (fast) {
            // Single-workgroup fast path: block 0 runs consecutive small
            // levels with workgroup barriers only (no grid tier: it alone
            // works); the rest of the grid waits at the next grid barrier.
            // Every workgroup must be done reading this step's counters
            // before block 0 rewrites them (its second fast step zeroes
            // qlen / fedges of index step % 3, which a workgroup still on
            // its way out of the previous barrier would otherwise read as an
            // empty frontier and leave the loop, stranding the grid at the
            // barrier below).  The per-level path rewrites counters as it
            // goes and takes an entry barrier.  The LDS path writes them
            // only when it ends, so it starts at once and, before writing,
            // waits for the others' arrival at the closing barrier -- except
            // after a folded search start, whose label rewrite the entry
            // barrier orders before the first claim.
            const bool lds_path = LOOP == kLoopEC || LOOP == kLoopDO || contract_lds;
            if (!lds_path || first) bar(!first);
            if (lds_path) {
                // frontier kept in LDS between levels
                if (blockIdx.x == 0) {
                    unsigned long long t0 = s.instrument ? wall_clock64() : 0;
                    fast_lds_path<PRED, EXACT, FASTV>(g, s, c, n, loop, vsum, fe, shared.fast, lds_path && !first,
                                                      bshape);
                    __syncthreads();
                    init_lds(lds);
                    if (s.instrument) busy += wall_clock64() - t0;
                }
            } else if (blockIdx.x == 0) {
                unsigned long long t0 = s.instrument ? wall_clock64() : 0;
                unsigned long long fv = vsum;  // sum through the step being run
                // (the step bodies grab every tile: see tile_grab)
                if (threadIdx.x == 0) lds.solo = 1;
                __syncthreads();
                while (true) {
                    if (threadIdx.x == 0) prepare_step(s.ctl, c, n);
                    loop_body<PRED, EXACT, LOOP>(c.phase, g, s, c, n, lds, 0, false);
                    // the next level reads queue entries other waves just
                    // stored: each wave's stores must have completed, which
                    // __syncthreads() alone does not wait for
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                    c = advance(loop, c);
                    if (threadIdx.x == 0) lds.nb = frontier_length(s, c.step);
                    __syncthreads();
                    n = lds.nb;
                    if (n == 0 || !fast_level(s, c, n) || ld_agent(&s.ctl->ovf[c.step % 3])) break;
                    if (loop == kLoopCE && s.sat_t >= 0 && n > (unsigned) s.sat_t) break;
                    // (a small two-phase contraction is the LDS path's entry)
                    if (loop == kLoopTP && c.phase == kPhaseContract && n <= (unsigned) kFastCap) break;
                    // (direction-optimizing searches always take the LDS path)
                }
                if (threadIdx.x == 0) {
                    st_agent(&s.ctl->fast_step, c.step);
                    st_agent(&s.ctl->fast_iteration, c.iteration);
                    st_agent(&s.ctl->fast_phase, c.phase);
                    // excludes the step the loop top runs next (it adds it)
                    __hip_atomic_store(&s.ctl->fast_vsum, fv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    lds.solo = 0;  // (the grid barrier below orders it)
                }
                if (s.instrument) busy += wall_clock64() - t0;
            }
            bar(true);
            c.step = ld_agent(&s.ctl->fast_step);
            c.iteration = ld_agent(&s.ctl->fast_iteration);
            c.phase = ld_agent(&s.ctl->fast_phase);
            vsum = ld64_agent(&s.ctl->fast_vsum);
            prev_fe = ld64_agent(&s.ctl->fast_prev_fe);
            continue;
        }

        unsigned long long t0 = s.instrument ? wall_clock64() : 0;
        // A top-down step of a direction-optimizing search marks its claims
        // in the next frontier's bitmap when that bitmap is clear, so a turn
        // to bottom-up after it needs no bitmap pass.
        unsigned *next_bits = nullptr;
        if (LOOP == kLoopDO && !mark && c.phase == kPhaseExpandContract &&
            !(fbm_dirty & (1u << ((c.iteration + 1) & 1)))) {
            next_bits = s.fbm[(c.iteration + 1) & 1];
            fbm_dirty |= 1u << ((c.iteration + 1) & 1);
            fbm_exact_step = c.step + 1;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) prepare_step(s.ctl, c, n);
        if (spread) {
            if constexpr (LOOP == kLoopEC || LOOP == kLoopDO) {
                spread_step<PRED, EXACT>(g, s, c, n, shared.fast.f, next_bits);
                __syncthreads();
                init_lds(lds);
            }
        } else if (mark) {
            body_expand_mark<PRED, EXACT>(g, s, c, n, lds, 0, true);
        } else {
            loop_body<PRED, EXACT, LOOP>(c.phase, g, s, c, n, lds, 0, true, next_bits, from_bitmap);
        }
        if (s.instrument) busy += wall_clock64() - t0;
        // (light: a top-down or contract/expand step of at most
        // kLightWork frontier edges, or items when edges are not tracked)
        const bool light_step = c.phase != kPhaseBottomUp && !mark &&
                                (s.track_edges ? fe : (unsigned long long) n) <= kLightWork;
        bar(light_step);
        if (!spread && gathers(c.phase) && hubs_pending(s, c.step)) {
            t0 = s.instrument ? wall_clock64() : 0;
            if (mark) body_expand_mark<PRED, EXACT>(g, s, c, n, lds, 1, true);
            else loop_body<PRED, EXACT, LOOP>(c.phase, g, s, c, n, lds, 1, true, next_bits, from_bitmap);
            if (s.instrument) busy += wall_clock64() - t0;
            bar(false);
        }
        if (mark) {
            t0 = s.instrument ? wall_clock64() : 0;
            unsigned *fbm_out = nullptr;
            if (LOOP == kLoopDO) {
                fbm_out = s.fbm[(c.iteration + 1) & 1];
                fbm_dirty |= 1u << ((c.iteration + 1) & 1);
                fbm_exact_step = c.step + 1;
            }
            mark_compact<PRED, EXACT>(g, s, c, lds, fbm_out);
            if (s.instrument) busy += wall_clock64() - t0;
            bar(false);
        }
        // (fbm_dirty kept this step's claims out of the bitmap it cleared)
        if (from_bitmap) fbm_dirty = 0u;
        c = advance(loop, c);
    }

    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s.ctl->exit_reason = reason;
        s.ctl->exit_step = c.step;
        s.ctl->exit_iteration = c.iteration;
        s.ctl->exit_phase = c.phase;
        if (s.host_head) publish_status(s);
    }
    if (s.instrument && threadIdx.x == 0) {
        unsigned long long life = wall_clock64() - t_begin;
        atomicAdd(&s.ctl->busy_cycles, busy);
        atomicAdd(&s.ctl->life_cycles, life);
    }
}

// ---------------------------------------------------------------------------
// Launches per step: the tile phase, then (for gathering phases) the hub
// phase as its own launch.
// ---------------------------------------------------------------------------
template <bool PRED, bool EXACT, int PHASE>
__global__ __launch_bounds__(kBlock) void step_kernel(GraphArgs g, SearchArgs s, Cursor c)
{
    __shared__ Lds<PRED> lds;
    if (s.hybrid_steps && ld_agent(&s.ctl->hybrid_stop)) return;
    unsigned n = frontier_length(s, c.step);
    if (ld_agent(&s.ctl->overflow)) n = 0;
    if (PHASE == kPhaseContract && s.hybrid_steps && n > 0 && n <= (unsigned) s.sat_t) {
        // The filtered edge frontier fits the persistent kernel again: it
        // runs this step (enactor_hybrid.cuh:386-532 leaves the two-phase
        // loop at the same point).  Every workgroup sees the same n.
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            s.ctl->exit_step = c.step;
            s.ctl->exit_iteration = c.iteration;
            st_agent(&s.ctl->hybrid_stop, 1u);
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) prepare_step(s.ctl, c, n);
    if (n == 0) return;
    init_lds(lds);
    run_body<PRED, EXACT>(PHASE, g, s, c, n, lds, 0, true);
}

template <bool PRED, bool EXACT, int PHASE>
__global__ __launch_bounds__(kBlock) void hub_kernel(GraphArgs g, SearchArgs s, Cursor c)
{
    __shared__ Lds<PRED> lds;
    if (s.hybrid_steps && ld_agent(&s.ctl->hybrid_stop)) return;
    if (!hubs_pending(s, c.step)) return;
    init_lds(lds);
    // the tile step's input length: it sets the chunk size of the list
    run_body<PRED, EXACT>(PHASE, g, s, c, frontier_length(s, c.step), lds, 1, true);
}

__global__ __launch_bounds__(kBlock) void frontier_bitmap_kernel(GraphArgs g, SearchArgs s, Cursor c, int pass)
{
    frontier_bitmap_pass(g, s, c, frontier_length(s, c.step), pass);
}

__global__ __launch_bounds__(kBlock) void queue_from_bitmap_kernel(GraphArgs g, SearchArgs s, Cursor c)
{
    __shared__ Lds<false> lds;
    queue_from_bitmap<false>(g, s, c, lds);
}

// ---------------------------------------------------------------------------
// Initialisation kernels.
// ---------------------------------------------------------------------------
__global__ void reset_labels_kernel(int *depth, int *pred, unsigned *visited, int nodes)
{
    int words = (nodes + 31) >> 5;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += gridDim.x * blockDim.x) {
        depth[i] = -1;
        if (pred) pred[i] = -1;
        if (i < words) visited[i] = 0u;
    }
}

// Seeds the search: counters to zero, the source in queue 0 (predecessor -2),
// and for vertex-frontier strategies the source already labelled.  Resetting
// every counter here, before the first step, replaces the reference's
// iteration-0 counter reset inside every CTA (contract_expand_atomic/
// kernel.cuh:307-332), which can race with CTA 0's first enqueue.
// Starts a search: clears the control block, queues the source and, with
// reset set, rewrites every label in the same pass (each element written by
// exactly one thread, the source's with its start values), saving a launch.
__global__ void init_search_kernel(SearchArgs s, const int *row, int src, int label_source, int reset)
{
    if (reset) {
        const int words = (s.nodes + 31) >> 5;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.nodes; i += gridDim.x * blockDim.x) {
            const bool is_src = label_source && i == src;
            s.depth[i] = is_src ? 0 : -1;
            if (s.pred) s.pred[i] = is_src ? -2 : -1;
            if (i < words) {
                const bool holds = label_source && (src >> 5) == i;
                s.visited[i] = holds ? 1u << (src & 31) : 0u;
            }
        }
    }
    if (blockIdx.x != 0) return;
    Control *ctl = s.ctl;
    unsigned *words = reinterpret_cast<unsigned *>(ctl);
    for (int i = threadIdx.x; i < (int) (sizeof(Control) / 4); i += blockDim.x) words[i] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->qlen[0] = 1;
        ctl->max_depth = -1;
        s.q[0][0] = src;
        if (s.pq[0]) s.pq[0][0] = -2;
        if (label_source) {
            if (!reset) {
                s.depth[src] = 0;
                if (s.pred) s.pred[src] = -2;
                s.visited[src >> 5] |= 1u << (src & 31);
            }
            ctl->max_depth = 0;
            unsigned long long deg = (unsigned long long) (row[src + 1] - row[src]);
            ctl->fedges[0] = deg;
            ctl->visited_edges = deg;
        }
    }
}

// ---------------------------------------------------------------------------
// Barrier microbenchmark and primitive self-test.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void barrier_bench_kernel(Control *ctl, int iterations)
{
    for (int i = 0; i < iterations; i++) grid_barrier(ctl);
}

}  // namespace b40c_hip

// Plain 16-byte-vector streaming copy (counter calibration).
extern "C" __global__ __launch_bounds__(256) void b40c_calibration_copy_kernel(const uint4 *src, uint4 *dst,
                                                                              size_t n)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
        dst[i] = src[i];
}

namespace b40c_hip {

__global__ __launch_bounds__(kWave) void wave_selftest_kernel(const int *in, int *out_scan, int *out_rank, int n)
{
    int i = blockIdx.x * kWave + threadIdx.x;
    int x = i < n ? in[i] : 0;
    int s = wave_inclusive_sum(x);
    uint64_t m = ballot((x & 1) != 0);
    int r = lanes_below(m);
    if (i < n) {
        out_scan[i] = s;
        out_rank[i] = r;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {

template <bool PRED, bool EXACT>
const void *persistent_ptr(int loop, bool fastv)
{
    switch (loop) {
    case kLoopEC:
        return fastv ? reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopEC, true>)
                     : reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopEC, false>);
    case kLoopCE:
        return fastv ? reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopCE, true>)
                     : reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopCE, false>);
    case kLoopTP:
        return fastv ? reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopTP, true>)
                     : reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopTP, false>);
    default:
        return fastv ? reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopDO, true>)
                     : reinterpret_cast<const void *>(&persistent_kernel<PRED, EXACT, kLoopDO, false>);
    }
}

// fastv: the instantiation whose fast path keeps the visited bitmap in LDS
const void *persistent_select(int loop, bool pred, bool exact, bool fastv)
{
    if (pred) return exact ? persistent_ptr<true, true>(loop, fastv) : persistent_ptr<true, false>(loop, fastv);
    return exact ? persistent_ptr<false, true>(loop, fastv) : persistent_ptr<false, false>(loop, fastv);
}

bool uses_fast_visited(const SearchArgs &s) { return s.fast_visited && s.nodes <= kFastVisitedWords * 32; }

template <bool PRED, bool EXACT, int PHASE>
hipError_t launch_phase(const GraphArgs &g, const SearchArgs &s, Cursor c, int grid, hipStream_t st, int *launches)
{
    hipLaunchKernelGGL((step_kernel<PRED, EXACT, PHASE>), dim3(grid), dim3(kBlock), 0, st, g, s, c);
    (*launches)++;
    if (PHASE == kPhaseExpandContract || PHASE == kPhaseContractExpand || PHASE == kPhaseExpand) {
        hipLaunchKernelGGL((hub_kernel<PRED, EXACT, PHASE>), dim3(grid), dim3(kBlock), 0, st, g, s, c);
        (*launches)++;
    }
    return hipGetLastError();
}

template <bool PRED, bool EXACT>
hipError_t launch_step_t(int phase, const GraphArgs &g, const SearchArgs &s, Cursor c, int grid, hipStream_t st,
                         int *launches)
{
    switch (phase) {
    case kPhaseExpandContract: return launch_phase<PRED, EXACT, kPhaseExpandContract>(g, s, c, grid, st, launches);
    case kPhaseContractExpand: return launch_phase<PRED, EXACT, kPhaseContractExpand>(g, s, c, grid, st, launches);
    case kPhaseContract: return launch_phase<PRED, EXACT, kPhaseContract>(g, s, c, grid, st, launches);
    case kPhaseExpand: return launch_phase<PRED, EXACT, kPhaseExpand>(g, s, c, grid, st, launches);
    case kPhaseFilter: return launch_phase<PRED, EXACT, kPhaseFilter>(g, s, c, grid, st, launches);
    case kPhaseBottomUp: return launch_phase<PRED, EXACT, kPhaseBottomUp>(g, s, c, grid, st, launches);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_reset_labels(int *depth, int *pred, unsigned *visited, int nodes, hipStream_t st)
{
    int grid = (nodes + 255) / 256;
    if (grid > 16384) grid = 16384;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(reset_labels_kernel, dim3(grid), dim3(256), 0, st, depth, pred, visited, nodes);
    return hipGetLastError();
}

hipError_t launch_init_search(const SearchArgs &s, const int *row, int src, int label_source, bool reset,
                              hipStream_t st)
{
    int grid = 1;
    if (reset) grid = std::min(16384, std::max(1, (s.nodes + 255) / 256));
    hipLaunchKernelGGL(init_search_kernel, dim3(grid), dim3(256), 0, st, s, row, src, label_source, reset ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_queue_from_bitmap(const SearchArgs &s, Cursor c, int grid, hipStream_t st, int *launches)
{
    GraphArgs g{nullptr, nullptr, s.nodes, nullptr, nullptr};
    hipLaunchKernelGGL(queue_from_bitmap_kernel, dim3(grid), dim3(kBlock), 0, st, g, s, c);
    (*launches)++;
    return hipGetLastError();
}

hipError_t launch_build_frontier_bitmap(const SearchArgs &s, Cursor c, int grid, hipStream_t st, int *launches)
{
    GraphArgs g{nullptr, nullptr, s.nodes, nullptr, nullptr};
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(frontier_bitmap_kernel, dim3(grid), dim3(kBlock), 0, st, g, s, c, pass);
        (*launches)++;
    }
    return hipGetLastError();
}

hipError_t launch_step(int phase, bool pred, bool exact, const GraphArgs &g, const SearchArgs &s, Cursor c,
                       int grid, hipStream_t st, int *launches)
{
    if (pred) return exact ? launch_step_t<true, true>(phase, g, s, c, grid, st, launches)
                           : launch_step_t<true, false>(phase, g, s, c, grid, st, launches);
    return exact ? launch_step_t<false, true>(phase, g, s, c, grid, st, launches)
                 : launch_step_t<false, false>(phase, g, s, c, grid, st, launches);
}

hipError_t launch_coresident(const void *kernel, dim3 grid, dim3 block, void **args, hipStream_t st)
{
    static std::atomic<bool> plain{false};
    if (!plain.load(std::memory_order_relaxed)) {
        hipError_t e = hipLaunchCooperativeKernel(kernel, grid, block, args, 0, st);
        if (e != hipErrorIllegalState) return e;
        (void) hipGetLastError();  // the refusal is not an error of this launch
        plain.store(true, std::memory_order_relaxed);
    }
    return hipLaunchKernel(kernel, grid, block, args, 0, st);
}

hipError_t launch_persistent(int loop, bool pred, bool exact, const GraphArgs &g, const SearchArgs &s, Cursor c,
                             int grid, hipStream_t st)
{
    GraphArgs ga = g;
    SearchArgs sa = s;
    Cursor ca = c;
    void *args[] = {&ga, &sa, &ca};
    return launch_coresident(persistent_select(loop, pred, exact, uses_fast_visited(s)), dim3(grid), dim3(kBlock),
                             args, st);
}

// The smaller of the two instantiations' occupancies (which one a search
// launches depends on the graph and the configuration).
int persistent_max_blocks_per_cu(int loop, bool pred, bool exact)
{
    int blocks = 1 << 30;
    for (int fastv = 0; fastv < 2; fastv++) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, persistent_select(loop, pred, exact, fastv != 0), kBlock,
                                                         0) != hipSuccess)
            return 0;
        blocks = std::min(blocks, b);
    }
    return blocks;
}

int step_max_blocks_per_cu(bool pred)
{
    int blocks = 0;
    const void *k = pred ? reinterpret_cast<const void *>(&step_kernel<true, true, kPhaseExpandContract>)
                         : reinterpret_cast<const void *>(&step_kernel<false, true, kPhaseExpandContract>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, kBlock, 0) != hipSuccess) return 0;
    return blocks;
}

int barrier_bench_max_blocks_per_cu()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void *>(&barrier_bench_kernel),
                                                     kBlock, 0) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_barrier_bench(Control *ctl, int iterations, int grid, hipStream_t st)
{
    Control *c = ctl;
    int it = iterations;
    void *args[] = {&c, &it};
    return launch_coresident(reinterpret_cast<const void *>(&barrier_bench_kernel), dim3(grid), dim3(kBlock), args,
                             st);
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    size_t n = bytes / 16;
    hipLaunchKernelGGL(b40c_calibration_copy_kernel, dim3(2048), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(src), reinterpret_cast<uint4 *>(dst), n);
    return hipGetLastError();
}

hipError_t launch_wave_selftest(const int *in, int *out_scan, int *out_rank, int n, hipStream_t st)
{
    int grid = (n + kWave - 1) / kWave;
    hipLaunchKernelGGL(wave_selftest_kernel, dim3(grid), dim3(kWave), 0, st, in, out_scan, out_rank, n);
    return hipGetLastError();
}

}  // namespace b40c_hip
