// bfs_device.hpp -- gfx950 device building blocks of the BFS kernels.
//
// Wave64 primitives (DPP prefix sum, ballot/mbcnt compaction), the per-wave
// LDS-staged frontier queue, the visited-bitmask claims, the three-tier
// neighbour gather and the hierarchical grid barrier.  These replace the
// reference's CTA raking scans, warp scans, tile I/O, culls and software
// global barrier (b40c/util/scan/*.cuh, b40c/util/io/*.cuh,
// b40c/util/global_barrier.cuh, b40c/graph/bfs/*/cta.cuh); the designs are
// written for 64-lane waves and CDNA memory scopes, not translated.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace b40c_hip {

constexpr int kWave = 64;
constexpr int kHistory = 1024;                 // LDS history-hash entries per workgroup
constexpr int kBarrierFanIn = 16;              // first-level barrier counters
constexpr int kMaxXcdGrid = 1024;              // largest grid whose barriers may use XCD leaves
constexpr int kHubChunk = 2048;                // edges per grid-tier chunk
constexpr int kLevelLog = 64;                  // steps with recorded length/clock

// ---------------------------------------------------------------------------
// Device control block: frontier lengths, tile counters, barrier state, exit
// status.  Counters rotate by step so a step never zeroes a counter that is
// still being read: step s reads qlen[s % 3], appends to qlen[(s+1) % 3] and
// zeroes qlen[(s+2) % 3] and steal[(s+1) & 1].
// ---------------------------------------------------------------------------
struct alignas(256) Control {
    unsigned qlen[4];
    unsigned steal[2];
    unsigned overflow;
    int exit_reason;          // 0 done, 1 saturated, 2 overflow
    int exit_step;            // loop position when a persistent launch ends
    int exit_iteration;
    int exit_phase;
    int fast_step;            // single-workgroup fast path hand-off
    int fast_iteration;
    int fast_phase;
    int max_depth;            // deepest label written
    unsigned compact;         // bitmap -> queue reservations (zeroed every step)
    unsigned long long total_queued;
    unsigned long long busy_cycles;   // instrumented: sum over workgroups of step time
    unsigned long long life_cycles;   // instrumented: sum over workgroups of kernel time
    unsigned long long hub_alloc[2];  // grid tier: (entries << 32) | chunks, by step & 1
    unsigned hub_steal[2];            // grid tier: chunk counter, by step & 1
    unsigned hybrid_stop;             // hybrid two-phase launches: the filtered frontier fits the
                                      // persistent kernel again (exit_step / exit_iteration say where)
    unsigned pad1[3];
    // Direction-optimizing bookkeeping (second 128-byte line).
    unsigned long long fedges[3];     // sum of degrees of each frontier, by step % 3
    unsigned long long visited_edges; // sum of degrees of all vertices labelled so far
    unsigned long long fast_vsum;     // fast-path hand-off of the per-workgroup visited-edge sum
    unsigned ovf[4];                  // per-step overflow of the output queue, by step % 3
    unsigned long long fast_prev_fe;  // fast-path hand-off: frontier edges of the level before the next step's
    unsigned long long pad2[8];
    unsigned bar_root;        // own 128-byte line
    unsigned pad3[31];
    unsigned bar_gen;         // completed barriers (own 128-byte line)
    unsigned pad4[31];
    unsigned bar_leaf[kBarrierFanIn * 32];     // one 128-byte line per leaf counter
    unsigned bar_release[kBarrierFanIn * 32];  // one 128-byte line per leaf release word
    unsigned level_len[kLevelLog];             // input length of each step
    unsigned long long level_clock[kLevelLog]; // wall clock at the start of each step
    unsigned xcd_of[kMaxXcdGrid];              // XCD of each workgroup of a persistent launch
};

// Bytes of the control block's status head (counters, flags, statistics):
// what the host reads after a search or a step.
constexpr size_t kControlHead = 256;
static_assert(offsetof(Control, bar_root) == kControlHead, "status head ends where the barrier lines begin");

// Grid-tier work item: a row too long for one workgroup, split into
// chunk-edge chunks [first_chunk, first_chunk + ceil(len / chunk)); the
// chunk is kHubChunk, or kBlock for steps with fewer tiles than workgroups.
struct HubEntry {
    int beg, end, par;
    unsigned first_chunk;
    int tag;   // owner of the row (the query, in batched searches)
};

enum ExitReason { kExitDone = 0, kExitSaturated = 1, kExitOverflow = 2 };

// ---------------------------------------------------------------------------
// Threads per workgroup, chosen per translation unit.  The single-source
// kernels run 512: a grid barrier's release and acquire fences cost about
// the same per workgroup, so fewer, larger workgroups halve its cost at the
// same wave count.  The batch kernels build with 256 (B40C_BLOCK=256).
// Everything below depends on it and sits in an inline namespace named
// after the size, so the two builds never share a symbol.
// ---------------------------------------------------------------------------
#ifndef B40C_BLOCK
#define B40C_BLOCK 512
#endif
#define B40C_NS_CAT2(a, b) a##b
#define B40C_NS_CAT(a, b) B40C_NS_CAT2(a, b)
inline namespace B40C_NS_CAT(block, B40C_BLOCK) {

constexpr int kBlock = B40C_BLOCK;
constexpr int kWaves = kBlock / kWave;         // waves per workgroup
static_assert(kBlock % kWave == 0 && kBlock >= 128 && kBlock <= 1024, "workgroup of whole waves");

// LDS staging entries per wave of the frontier queues.  Every full area
// costs one reservation atomic on the queue's length, and on a large
// level those atomics, all on one address, bound the step: 2048 entries
// per wave took the single-source two-phase search on R-MAT s22 from 5.09
// to 2.64 ms (64 KiB of LDS per 512-thread workgroup, 128 KiB with the
// predecessor staging).  The batch kernels keep 512, which leaves them
// their occupancy of three 256-thread workgroups per CU.
constexpr int kWaveQueueCap = kBlock >= 512 ? 2048 : 512;

// ---------------------------------------------------------------------------
// Wave64 primitives
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id()
{
    return (int) __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Number of set bits of `m` below the calling lane.
__device__ __forceinline__ int lanes_below(uint64_t m)
{
    return (int) __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32),
                                           __builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int readlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

// Inclusive prefix sum across the 64 lanes with DPP row shifts inside each
// 16-lane row followed by the row_bcast:15 / row_bcast:31 carries between
// rows (GFX9/CDNA DPP controls 0x111..0x118, 0x142, 0x143).  Lanes whose
// DPP source is out of the row, or whose row is masked off, add 0.
__device__ __forceinline__ int wave_inclusive_sum(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
    return x;
}

// Inclusive running maximum of non-negative values across the 64 lanes (the
// same DPP steps; lanes with no source take 0, the identity here).  The last
// lane holds the wave's maximum.
__device__ __forceinline__ int wave_inclusive_max_nonneg(int x)
{
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false));
    return x;
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

// ---------------------------------------------------------------------------
// Scoped memory helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned ld_agent(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Queue entries another workgroup stored in an earlier step, read once:
// system scope makes the load miss in every cache level, so a line this XCD
// still holds from an older use of the buffer (an earlier step or search)
// cannot be returned.  Streamed data, so nothing is lost by not caching it.
__device__ __forceinline__ int ld_queue(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_agent(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned add_agent(unsigned *p, unsigned v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Visited claims.
//
// Exact: the bitmask word is first read with a plain (cacheable, possibly
// stale) load.  Bits only ever go 0 -> 1 during a search, so a set bit is
// always true and skips the atomic; a clear bit may be stale and is settled
// by the agent-scope atomic OR, whose returned old value makes exactly one
// claimant the owner of the vertex.
//
// Best effort (the reference discipline, contract_expand_atomic/cta.cuh:
// 213-304): racy read-modify-write of the mask byte as a pre-filter, then the
// label check; racing discoverers may both succeed, which only duplicates
// work because they write the same depth.
// ---------------------------------------------------------------------------
// The pre-test read decides only whether to try the atomic.  The cached read
// may miss bits other workgroups set during this level (device-scope atomics
// are performed beyond this XCD's L2, which keeps its old copy); the
// coherent read sees them at the cost of a longer trip.
__device__ __forceinline__ bool claim_exact(unsigned *visited, int u, bool coherent = false)
{
    unsigned *w = visited + (u >> 5);
    unsigned bit = 1u << (u & 31);
    unsigned cur = coherent ? __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *w;
    if (cur & bit) return false;
    unsigned old = __hip_atomic_fetch_or(w, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (old & bit) == 0;
}

__device__ __forceinline__ bool claim_best_effort(unsigned *visited, const int *depth, int u)
{
    unsigned *w = visited + (u >> 5);
    unsigned bit = 1u << (u & 31);
    unsigned cur = *w;
    if (cur & bit) return false;
    *w = cur | bit;
    return depth[u] == -1;
}

template <bool EXACT>
__device__ __forceinline__ bool claim(unsigned *visited, const int *depth, int u, bool coherent = false)
{
    if (EXACT) return claim_exact(visited, u, coherent);
    return claim_best_effort(visited, depth, u);
}

// ---------------------------------------------------------------------------
// LDS layout shared by every BFS kernel.
// ---------------------------------------------------------------------------
template <bool PRED>
struct PredStore {
    int buf[kWaves][kWaveQueueCap];
};
template <>
struct PredStore<false> {
    int buf[1][1];
};

template <bool PRED>
struct Lds {
    int wq[kWaves][kWaveQueueCap];       // per-wave frontier staging
    PredStore<PRED> wqp;                 // per-wave predecessor staging
    int sc_excl[kWaves][kWave];          // scan tier: exclusive rank per lane
    int sc_beg[kWaves][kWave];           // scan tier: row begin per lane
    int sc_par[kWaves][kWave];           // scan tier: row owner per lane
    int hist[kHistory];                  // history hash (contract / filter)
    int cta_owner, cta_beg, cta_end, cta_par;
    unsigned tile;
    unsigned nb;                         // broadcast frontier length
    int step, iteration;
    // end-of-step reduction over the waves (one global update per workgroup)
    unsigned ep_count[kWaves];
    unsigned ep_extra[kWaves];
    unsigned long long ep_fe[kWaves];
    int ep_lab[kWaves];
    int ep_scan[kWaves];
    unsigned ep_base;
    int solo;                            // 1: this workgroup runs the step alone (no static tiles)
};

// ---------------------------------------------------------------------------
// Per-wave frontier queue.  Lanes append through ballot + mbcnt into the
// wave's LDS staging area; a full area (or the end of a step) is flushed with
// one agent-scope atomic reservation and coalesced stores.  One global atomic
// per kWaveQueueCap appended elements (or per wave per step) replaces the
// reference's per-tile raking scan + atomicAdd (util/scan/cooperative_scan.cuh:
// 172-195, 309-343).
// ---------------------------------------------------------------------------
template <bool PRED>
struct WaveQueue {
    int *buf;
    int *pbuf;
    int count;          // wave-uniform
    int *out;
    int *pout;
    unsigned *out_len;
    unsigned cap;
    unsigned *overflow;
    unsigned *overflow_step;   // per-step flag the level loop decides on (may alias overflow)

    __device__ __forceinline__ void flush()
    {
        if (count == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        int lane = lane_id();
        unsigned base = 0;
        if (lane == 0) base = add_agent(out_len, (unsigned) count);
        base = (unsigned) uniform((int) base);
        unsigned lim = 0;
        if (base < cap) lim = min((unsigned) count, cap - base);
        if (lim < (unsigned) count && lane == 0) {
            st_agent(overflow, 1u);
            st_agent(overflow_step, 1u);
        }
        for (unsigned i = lane; i < lim; i += kWave) {
            out[base + i] = buf[i];
            if (PRED) pout[base + i] = pbuf[i];
        }
        count = 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }

    // Wave-converged: every lane of the wave calls push together.
    __device__ __forceinline__ void push(bool flag, int u, int p)
    {
        uint64_t m = ballot(flag);
        if (m == 0) return;
        int c = __popcll(m);
        if (count + c > kWaveQueueCap) flush();
        if (flag) {
            int pos = count + lanes_below(m);
            buf[pos] = u;
            if (PRED) pbuf[pos] = p;
        }
        count += c;
    }
};

// ---------------------------------------------------------------------------
// Three-tier neighbour gather.  Each thread owns one adjacency row
// [beg, end) with owner `par` (end == beg for none).  fn(valid, u, par) is
// called wave-converged for every neighbour u exactly once:
//   CTA tier  (len >= cta_t):  the whole workgroup strides through the row;
//   wave tier (len >= wave_t): one wave strides through the row;
//   scan tier (len <  wave_t): a wave-wide DPP prefix sum ranks the short
//       rows, and each lane locates its edge by a 6-step search over the 64
//       LDS-staged exclusive ranks, so the wave's lanes stay fully occupied
//       whatever the mix of short lengths.
// Replaces ExpandByCta / ExpandByWarp / ExpandByScan
// (b40c/graph/bfs/two_phase/expand_atomic/cta.cuh:215-423).
// ---------------------------------------------------------------------------
struct HubList {
    HubEntry *entries;
    unsigned cap;
    unsigned long long *alloc;
    int grid_t;        // rows at least this long go to the grid tier
    int tag;           // stored with every published row
    int chunk;         // edges per chunk
};

// Index of the last grid-tier entry whose first chunk is <= ch (entries are
// in chunk order).  A workgroup takes chunks in increasing order, so the
// search gallops forward from the entry of its previous chunk (`hint`): the
// next chunk usually lies in the same entry or the next one, one or two
// loads instead of a binary search's chain of dependent loads.
__device__ __forceinline__ unsigned hub_entry_of(const HubEntry *list, unsigned entries, unsigned ch, unsigned hint)
{
    unsigned lo = 0, hi = entries - 1;
    if (hint < entries && list[hint].first_chunk <= ch) {
        lo = hint;
        unsigned step = 1;
        while (lo + step < entries && list[lo + step].first_chunk <= ch) {
            lo += step;
            step <<= 1;
        }
        hi = min(lo + step, entries) - 1;
    }
    while (lo < hi) {
        const unsigned mid = (lo + hi + 1) >> 1;
        if (list[mid].first_chunk <= ch) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Grid-tier publication of a row of at least hub.grid_t neighbours (called
// wave-converged; true when this lane's row was published).  The wave
// reserves list entries and chunk numbers for all its long rows with one
// atomic; entries keep lane order, so the list stays sorted by first chunk.
__device__ __forceinline__ bool publish_hub_row(int beg, int end, int par, const HubList &hub)
{
    const int lane = lane_id();
    const int len = end - beg;
    const bool pub = len >= hub.grid_t;
    const uint64_t pm = ballot(pub);
    if (!pm) return false;
    const int nch = pub ? (len + hub.chunk - 1) / hub.chunk : 0;
    const int incl = wave_inclusive_sum(nch);
    const unsigned total = (unsigned) readlane(incl, kWave - 1);
    unsigned lo = 0, hi = 0;
    if (lane == 0) {
        unsigned long long old = __hip_atomic_fetch_add(hub.alloc, ((unsigned long long) __popcll(pm) << 32) | total,
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lo = (unsigned) old;
        hi = (unsigned) (old >> 32);
    }
    lo = (unsigned) uniform((int) lo);
    hi = (unsigned) uniform((int) hi);
    const unsigned slot = hi + (unsigned) lanes_below(pm);
    if (pub && slot < hub.cap) {
        hub.entries[slot] = HubEntry{beg, end, par, lo + (unsigned) (incl - nch), hub.tag};
        return true;
    }
    return false;
}

template <bool PRED, typename Fn>
__device__ __forceinline__ void gather_rows(int beg, int end, int par, const int *__restrict__ col,
                                            int wave_t, int cta_t, const HubList &hub, Lds<PRED> &lds, Fn &&fn)
{
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    int len = end - beg;

    // Grid tier: publish the row as chunks for every workgroup of the grid
    // (processed after this tile phase).  A full list falls back to the CTA
    // tier below.
    if (publish_hub_row(beg, end, par, hub)) len = 0;

    // CTA tier
    while (true) {
        __syncthreads();
        if (len >= cta_t) lds.cta_owner = tid;
        __syncthreads();
        int owner = lds.cta_owner;
        if (owner < 0) break;
        if (tid == owner) {
            lds.cta_beg = beg;
            lds.cta_end = end;
            lds.cta_par = par;
            len = 0;
        }
        __syncthreads();
        int b = lds.cta_beg, e = lds.cta_end, pv = lds.cta_par;
        __syncthreads();
        if (tid == 0) lds.cta_owner = -1;
        for (int k = b; k < e; k += kBlock) {
            int idx = k + tid;
            bool ok = idx < e;
            int u = ok ? __builtin_nontemporal_load(col + idx) : -1;
            fn(ok, u, pv);
        }
    }

    // Wave tier
    uint64_t wm = ballot(len >= wave_t);
    while (wm) {
        int l = uniform(__builtin_ctzll(wm));
        wm &= wm - 1;
        int b = readlane(beg, l), e = readlane(end, l), pv = readlane(par, l);
        for (int k = b; k < e; k += kWave) {
            int idx = k + lane;
            bool ok = idx < e;
            int u = ok ? __builtin_nontemporal_load(col + idx) : -1;
            fn(ok, u, pv);
        }
    }
    if (len >= wave_t) len = 0;

    // Scan tier
    int incl = wave_inclusive_sum(len);
    int total = readlane(incl, kWave - 1);
    if (total == 0) return;
    int *excl = lds.sc_excl[wave];
    int *rb = lds.sc_beg[wave];
    int *rp = lds.sc_par[wave];
    excl[lane] = incl - len;
    rb[lane] = beg;
    rp[lane] = par;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int r = 0; r < total; r += kWave) {
        int idx = r + lane;
        bool ok = idx < total;
        int u = -1, pv = -1;
        if (ok) {
            int l = 0;
#pragma unroll
            for (int s = 32; s > 0; s >>= 1)
                if (excl[l + s] <= idx) l += s;
            u = __builtin_nontemporal_load(col + rb[l] + (idx - excl[l]));
            pv = rp[l];
        }
        fn(ok, u, pv);
    }
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Grid barrier for cooperatively launched (co-resident) grids.
// Two-level arrival: workgroup b arrives on leaf counter b % kBarrierFanIn;
// the last arrival at a leaf arrives at the root; the last root arrival
// resets the counters and publishes the new generation into every leaf's
// release word, so waiting workgroups poll kBarrierFanIn different lines
// instead of one.
//
// Memory ordering costs one agent-scope release fence (L2 write-back) and one
// acquire fence (L1/L2 invalidate) per workgroup, issued by thread 0 around
// the workgroup barriers.  Polling uses relaxed loads: an acquire load inside
// the spin would invalidate the caches on every poll and slow every other
// workgroup still working.  The last arrivers chain acquire/release fences so
// every workgroup's pre-barrier writes happen-before every post-barrier read.
// Replaces util::GlobalBarrier (b40c/util/global_barrier.cuh:80-126).
// ---------------------------------------------------------------------------
// Which leaf a workgroup arrives on, how many workgroups share it, and how
// many leaves there are.  By block (b % kBarrierFanIn), every workgroup
// writes back its XCD's L2 before arriving.  With XCD leaves (shared), the
// leaf is the workgroup's XCD and only the leaf's last arriver writes back
// that XCD's L2 -- once for all of them: the L1 is write-through, so a
// workgroup's stores are in its XCD's L2 once they have completed, which
// every workgroup waits for before it arrives.
struct BarrierShape {
    unsigned leaf, leaf_size, leaves, shared;
};
static_assert(kBarrierFanIn >= 8, "an XCD leaf indexes the leaf words");

__device__ __forceinline__ BarrierShape barrier_shape_by_block()
{
    const unsigned nblocks = gridDim.x;
    BarrierShape b;
    b.leaf = blockIdx.x % kBarrierFanIn;
    b.leaves = nblocks < (unsigned) kBarrierFanIn ? nblocks : (unsigned) kBarrierFanIn;
    // workgroups on this leaf: ceil((nblocks - leaf) / fan_in)
    b.leaf_size = (nblocks - b.leaf + kBarrierFanIn - 1) / kBarrierFanIn;
    b.shared = 0u;
    return b;
}

// The XCD (accelerator complex die) the calling wave runs on.
__device__ __forceinline__ unsigned xcc_id()
{
    unsigned x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x & 7u;
}

__device__ __forceinline__ void grid_barrier(Control *ctl, const BarrierShape &shape)
{
    // Every wave drains its own outstanding memory operations first:
    // __syncthreads() alone does not wait for other waves' stores, and the
    // single write-back below (thread 0's release fence) must find all of
    // this workgroup's stores already in L2, or a store still in flight
    // stays dirty in this XCD's L2 and readers elsewhere see old memory.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned leaf = shape.leaf, leaves = shape.leaves, leaf_size = shape.leaf_size;
        unsigned *release = &ctl->bar_release[leaf * 32];
        // The generation comes from the one global word, which only the last
        // arriver of a barrier advances (after every workgroup has read it
        // here); waiters then wait for their leaf word to reach gen + 1.
        // Leaf words of leaves a smaller earlier grid never used may lag:
        // they are compared with gen + 1, never with each other.
        const unsigned gen = __hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // (with XCD leaves the leaf's last arriver writes the L2 back; the
        // generation must still be read before this arrival)
        if (shape.shared) __builtin_amdgcn_s_waitcnt(0);
        else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned *lc = &ctl->bar_leaf[leaf * 32];
        unsigned arrived = __hip_atomic_fetch_add(lc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool last = false;
        if (arrived == leaf_size - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_store(lc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // Relaxed stores to different lines may land in any order, and a
            // relaxed RMW does not wait for them: the reset must be visible
            // before the root arrival that can release this leaf's
            // workgroups, or their next arrival could be overwritten by it.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            unsigned r = __hip_atomic_fetch_add(&ctl->bar_root, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (r == leaves - 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                __hip_atomic_store(&ctl->bar_root, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&ctl->bar_gen, gen + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                // the new generation and the root reset before any release
                // word: a released workgroup reads bar_gen at its next
                // barrier, and an old value there would let it straight
                // through
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                for (unsigned l = 0; l < leaves; l++)
                    __hip_atomic_store(&ctl->bar_release[l * 32], gen + 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                last = true;
            }
        }
        if (!last) {
            while (__hip_atomic_load(release, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gen + 1)
                __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

__device__ __forceinline__ void grid_barrier(Control *ctl) { grid_barrier(ctl, barrier_shape_by_block()); }

// Thread 0 of block 0 only (shape: block 0's): waits until every other
// workgroup has arrived at the grid barrier in progress (block 0 itself has
// not), i.e. until every other workgroup is done with whatever it did
// before that barrier.  Block 0's leaf then holds all its workgroups but
// block 0 and the root all leaves but that one; neither count can move on
// before block 0 arrives, so the wait ends exactly once they are reached.
__device__ __forceinline__ void wait_other_arrivals(Control *ctl, const BarrierShape &shape)
{
    while (__hip_atomic_load(&ctl->bar_root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != shape.leaves - 1 ||
           __hip_atomic_load(&ctl->bar_leaf[shape.leaf * 32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
               shape.leaf_size - 1)
        __builtin_amdgcn_s_sleep(2);
}

// Thread 0 of every workgroup of a persistent launch (gridDim.x <=
// kMaxXcdGrid), before its first barrier: records the workgroup's XCD.
__device__ __forceinline__ void xcd_register(Control *ctl) { ctl->xcd_of[blockIdx.x] = xcc_id(); }

// Thread 0, after the launch's first barrier (every workgroup recorded its
// XCD before arriving at it): XCD leaves when all eight XCDs hold
// workgroups of this launch, else leaves by block.  The counts are kept
// in 16-bit fields, four XCDs to a word.
__device__ __forceinline__ BarrierShape xcd_shape(Control *ctl)
{
    const unsigned x = xcc_id();
    unsigned long long lo = 0, hi = 0;
    const uint4 *rec = reinterpret_cast<const uint4 *>(ctl->xcd_of);
    const unsigned words = gridDim.x / 4, tail = gridDim.x % 4;
    for (unsigned i = 0; i < words; i++) {
        const uint4 q = rec[i];
        const unsigned v[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const unsigned y = v[k] & 7u;
            if (y < 4) lo += 1ull << (16 * y);
            else hi += 1ull << (16 * (y - 4));
        }
    }
    for (unsigned i = 4 * words; i < 4 * words + tail; i++) {
        const unsigned y = ctl->xcd_of[i] & 7u;
        if (y < 4) lo += 1ull << (16 * y);
        else hi += 1ull << (16 * (y - 4));
    }
    bool all = true;
    unsigned mine = 0;
#pragma unroll
    for (unsigned k = 0; k < 8; k++) {
        const unsigned c = (unsigned) (((k < 4 ? lo : hi) >> (16 * (k & 3))) & 0xFFFFull);
        all = all && c > 0;
        if (k == x) mine = c;
    }
    if (!all) return barrier_shape_by_block();
    BarrierShape b;
    b.leaf = x;
    b.leaf_size = mine;
    b.leaves = 8u;
    b.shared = 1u;
    return b;
}

}  // inline namespace (workgroup size)
}  // namespace b40c_hip
