// b40c_primitives.hpp -- device building blocks of the single-pass scan,
// stream compaction and radix sort kernels (b40c_scan.hip,
// b40c_csr_build.hip): wave64 and workgroup sums over uint32, and the
// decoupled look-back that chains tile prefixes through one status word
// per tile.
//
// Status word of a tile (64 bits, written and read whole, so the value needs
// no fence of its own): bits 63..34 the launch epoch, bits 33..32 the state
// (0 not yet published, 1 the tile's own aggregate, 2 the inclusive prefix
// of every tile up to and including it), bits 31..0 the value.  A word from
// an older launch (another epoch) reads as not published, so the status
// array needs no clearing between launches.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b40c_hip {
namespace prim {

constexpr int kWave = 64;
constexpr unsigned kStateAggregate = 1u, kStatePrefix = 2u;

__device__ __forceinline__ int lane()
{
    return (int) __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ int lanes_below(uint64_t m)
{
    return (int) __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
}

// Inclusive sum across the 64 lanes, modulo 2^32 (DPP row shifts within each
// 16-lane row, then the row_bcast:15 / row_bcast:31 carries between rows).
__device__ __forceinline__ unsigned wave_inclusive_sum_u32(unsigned x)
{
    int v = (int) x;
    v = (int) ((unsigned) v + (unsigned) __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false));
    v = (int) ((unsigned) v + (unsigned) __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false));
    v = (int) ((unsigned) v + (unsigned) __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false));
    v = (int) ((unsigned) v + (unsigned) __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false));
    v = (int) ((unsigned) v + (unsigned) __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false));
    v = (int) ((unsigned) v + (unsigned) __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false));
    return (unsigned) v;
}

__device__ __forceinline__ unsigned wave_total_u32(unsigned x)
{
    return (unsigned) __builtin_amdgcn_readlane((int) wave_inclusive_sum_u32(x), kWave - 1);
}

// Exclusive sum over a workgroup of THREADS threads (one value each); the
// workgroup total goes to *total.  Uses wave_tot[THREADS / 64] of LDS.
template <int THREADS>
__device__ __forceinline__ unsigned block_exclusive_sum_u32(unsigned x, unsigned *wave_tot, unsigned *total)
{
    constexpr int kWaves = THREADS / kWave;
    const int w = threadIdx.x / kWave;
    const unsigned incl = wave_inclusive_sum_u32(x);
    if (lane() == kWave - 1) wave_tot[w] = incl;
    __syncthreads();
    unsigned off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kWaves; k++) {
        const unsigned t = wave_tot[k];
        if (k < w) off += t;
        tot += t;
    }
    *total = tot;
    __syncthreads();
    return off + incl - x;
}

__device__ __forceinline__ unsigned long long status_word(unsigned epoch, unsigned state, unsigned value)
{
    return ((unsigned long long) (((epoch & 0x3FFFFFFFu) << 2) | state) << 32) | value;
}

__device__ __forceinline__ void publish(unsigned long long *status, long long tile, unsigned epoch, unsigned state,
                                        unsigned value)
{
    __hip_atomic_store(status + tile, status_word(epoch, state, value), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of tile `tile` > 0: the sum of every earlier tile's
// aggregate.  Called by all 64 lanes of one wave; returns the prefix to each.
// Each probe reads the 64 newest unexamined predecessors at once (lane 0 the
// newest); each lane waits until its tile has published, then the wave adds
// the aggregates down to the newest inclusive prefix, or all 64 and probes
// further back.  Tiles before tile 0 read as a prefix of 0.  Predecessors
// publish their aggregate before they look back themselves, so the waits
// always end.
__device__ __forceinline__ unsigned lookback_prefix(const unsigned long long *status, long long tile, unsigned epoch)
{
    const int l = lane();
    unsigned prefix = 0;
    long long newest = tile - 1;
    while (true) {
        const long long t = newest - l;
        unsigned state = kStatePrefix, value = 0;
        if (t >= 0) {
            while (true) {
                const unsigned long long w =
                    __hip_atomic_load(status + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned hi = (unsigned) (w >> 32);
                if ((hi >> 2) == (epoch & 0x3FFFFFFFu) && (hi & 3u) != 0u) {
                    state = hi & 3u;
                    value = (unsigned) w;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        const uint64_t pm = __ballot(state == kStatePrefix);
        if (pm) {
            const int stop = __builtin_ctzll(pm);  // newest tile holding an inclusive prefix
            prefix += wave_total_u32(l <= stop ? value : 0u);
            return prefix;
        }
        prefix += wave_total_u32(value);
        newest -= kWave;
    }
}

}  // namespace prim
}  // namespace b40c_hip
