// b40c_scan.hip -- single-pass prefix scan and stable stream compaction
// (C ABI in include/b40c_scan.h).
//
// One kernel per call.  A workgroup takes the next tile index from an atomic
// counter, reduces its 4096 items, publishes the aggregate, looks back over
// the predecessors' status words for its exclusive prefix
// (prim::lookback_prefix), publishes its inclusive prefix and writes its
// outputs.  This replaces the reference's upsweep / spine / downsweep passes
// (b40c/scan/enactor.cuh:426-454, b40c/copy_if/enactor.cuh:436-461) and their
// spine storage.
#include "b40c_scan.h"
#include "b40c_primitives.hpp"
#include "b40c_scan_internal.hpp"
#include "bfs_internal.hpp"

#include <algorithm>
#include <mutex>

using namespace b40c_hip;

namespace b40c_hip {

namespace {

constexpr int kThreads = kScanThreads;
constexpr int kItems = kScanItems;

// Items of one thread: kItems consecutive elements from `base` (blocked
// arrangement), `fill` past the end.
__device__ __forceinline__ void load_blocked(const uint32_t *in, long long n, long long base, bool vec,
                                             uint32_t fill, uint32_t (&v)[kItems])
{
    if (vec && base + kItems <= n) {
        const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (int k = 0; k < kItems / 4; k++) {
            const uint4 q = p[k];
            v[4 * k] = q.x;
            v[4 * k + 1] = q.y;
            v[4 * k + 2] = q.z;
            v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; k++) v[k] = base + k < n ? in[base + k] : fill;
    }
}

__device__ __forceinline__ bool selected(uint32_t key, int op, uint32_t comparand)
{
    switch (op) {
    case B40C_SELECT_LESS: return key < comparand;
    case B40C_SELECT_NOT_EQUAL: return key != comparand;
    default: return (int32_t) key >= 0;
    }
}

// Tile index from the launch-order counter, and the tile's exclusive prefix
// given its aggregate (every thread calls it; the result is returned to all).
__device__ __forceinline__ long long take_tile(const ChainArgs &c, long long *s_tile)
{
    if (threadIdx.x == 0)
        *s_tile = (long long) (atomicAdd(c.tile_counter, 1ull) - c.tile_base);
    __syncthreads();
    return *s_tile;
}

__device__ __forceinline__ unsigned chain_prefix(const ChainArgs &c, long long tile, unsigned aggregate,
                                                 unsigned *s_prefix)
{
    if (threadIdx.x < prim::kWave) {
        unsigned prefix = 0;
        if (tile == 0) {
            if (threadIdx.x == 0) prim::publish(c.status, tile, c.epoch, prim::kStatePrefix, aggregate);
        } else {
            if (threadIdx.x == 0) prim::publish(c.status, tile, c.epoch, prim::kStateAggregate, aggregate);
            prefix = prim::lookback_prefix(c.status, tile, c.epoch);
            if (threadIdx.x == 0) prim::publish(c.status, tile, c.epoch, prim::kStatePrefix, prefix + aggregate);
        }
        if (threadIdx.x == 0) *s_prefix = prefix;
    }
    __syncthreads();
    return *s_prefix;
}

__global__ __launch_bounds__(kThreads) void scan_kernel(const uint32_t *in, uint32_t *out, long long n, int exclusive,
                                                        int vec, ChainArgs c, uint32_t *d_total)
{
    __shared__ long long s_tile;
    __shared__ unsigned s_wave[kThreads / prim::kWave], s_prefix;
    const long long tile = take_tile(c, &s_tile);
    const long long base = tile * kScanTile + (long long) threadIdx.x * kItems;
    uint32_t v[kItems];
    load_blocked(in, n, base, vec != 0, 0u, v);
    unsigned sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) sum += v[k];
    unsigned agg;
    const unsigned before = prim::block_exclusive_sum_u32<kThreads>(sum, s_wave, &agg);
    const unsigned prefix = chain_prefix(c, tile, agg, &s_prefix);
    unsigned run = prefix + before;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const unsigned x = v[k];
        v[k] = exclusive ? run : run + x;
        run += x;
    }
    if (vec && base + kItems <= n) {
        uint4 *p = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
        for (int k = 0; k < kItems / 4; k++) p[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kItems; k++)
            if (base + k < n) out[base + k] = v[k];
    }
    if (d_total && tile == c.tiles - 1 && threadIdx.x == 0) *d_total = prefix + agg;
}

__global__ __launch_bounds__(kThreads) void copy_if_kernel(const uint32_t *in, uint32_t *out, long long n, int op,
                                                           uint32_t comparand, int vec, ChainArgs c,
                                                           uint32_t *d_count)
{
    __shared__ long long s_tile;
    __shared__ unsigned s_wave[kThreads / prim::kWave], s_prefix;
    const long long tile = take_tile(c, &s_tile);
    const long long base = tile * kScanTile + (long long) threadIdx.x * kItems;
    uint32_t v[kItems];
    load_blocked(in, n, base, vec != 0, 0u, v);
    unsigned keep = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++)
        if (base + k < n && selected(v[k], op, comparand)) keep |= 1u << k;
    unsigned agg;
    const unsigned before = prim::block_exclusive_sum_u32<kThreads>((unsigned) __popc(keep), s_wave, &agg);
    const unsigned prefix = chain_prefix(c, tile, agg, &s_prefix);
    unsigned at = prefix + before;
#pragma unroll
    for (int k = 0; k < kItems; k++)
        if ((keep >> k) & 1u) out[at++] = v[k];
    if (d_count && tile == c.tiles - 1 && threadIdx.x == 0) *d_count = prefix + agg;
}

__global__ void fill_u32_kernel(uint32_t *p, long long n, uint32_t value)
{
    for (long long i = blockIdx.x * (long long) blockDim.x + threadIdx.x; i < n; i += (long long) gridDim.x * blockDim.x)
        p[i] = value;
}

}  // namespace

hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, long long n, bool exclusive, const ChainArgs &c,
                           uint32_t *d_total, hipStream_t st)
{
    const int vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    hipLaunchKernelGGL(scan_kernel, dim3((unsigned) c.tiles), dim3(kThreads), 0, st, in, out, n, exclusive ? 1 : 0,
                       vec, c, d_total);
    return hipGetLastError();
}

hipError_t launch_copy_if_u32(const uint32_t *in, uint32_t *out, long long n, int op, uint32_t comparand,
                              const ChainArgs &c, uint32_t *d_count, hipStream_t st)
{
    const int vec = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    hipLaunchKernelGGL(copy_if_kernel, dim3((unsigned) c.tiles), dim3(kThreads), 0, st, in, out, n, op, comparand,
                       vec, c, d_count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free scan (b40c_scan_noalloc): the tile counter and the status
// words live in one file-scope device array per GPU.  Word 0 is the counter;
// the status words start at word kNoAllocStatus (a 256-byte line of their
// own).  Every call clears what it uses with a fill kernel (no memset node
// if the caller captures the call in a graph) and uses epoch 0.
// ---------------------------------------------------------------------------
namespace {

constexpr long long kNoAllocMaxTiles = 16384;  // 2^26 items
constexpr int kNoAllocStatus = 32;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) unsigned long long scratch[kNoAllocStatus + kNoAllocMaxTiles];

__global__ __launch_bounds__(kThreads) void noalloc_clear_kernel(unsigned long long *words, long long tiles)
{
    for (long long i = blockIdx.x * (long long) kThreads + threadIdx.x; i < tiles; i += (long long) gridDim.x * kThreads)
        words[kNoAllocStatus + i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) words[0] = 0ull;
}

std::mutex g_noalloc_mu;
unsigned long long *g_noalloc_ptr[kNoAllocDevices];  // device address of `scratch`, per device

// The device address of this device's `scratch` (the array's name is a host
// stand-in, not a device address), looked up once per device.
hipError_t noalloc_scratch(int device, unsigned long long **out)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_noalloc_ptr[device] = static_cast<unsigned long long *>(ptr);
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

}  // namespace

// ---------------------------------------------------------------------------
// Chain state: the tile counter and the per-tile status words.
// ---------------------------------------------------------------------------
hipError_t ChainState::init(int dev)
{
    device = dev;
    hipError_t e = hipMalloc(&d_counter, sizeof(unsigned long long) + 4 * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    d_word = reinterpret_cast<uint32_t *>(d_counter + 1);
    return hipMemset(d_counter, 0, sizeof(unsigned long long) + 4 * sizeof(uint32_t));
}

hipError_t ChainState::prepare(long long items, int tile_items, ChainArgs *c, hipStream_t st)
{
    const long long tiles = std::max(1LL, (items + tile_items - 1) / tile_items);
    if (tiles > capacity) {
        // earlier launches on the stream may still read the old words
        if (d_status) {
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            (void) hipFree(d_status);
        }
        d_status = nullptr;
        capacity = 0;
        hipError_t e = hipMalloc(&d_status, sizeof(unsigned long long) * (size_t) tiles);
        if (e != hipSuccess) return e;
        // state 0 (unpublished) under every epoch; ordered on the launch's
        // (non-blocking) stream
        if ((e = hipMemsetAsync(d_status, 0, sizeof(unsigned long long) * (size_t) tiles, st)) != hipSuccess)
            return e;
        capacity = tiles;
    }
    epoch = (epoch + 1) & 0x3FFFFFFFu;
    if (epoch == 0) {
        // the 30-bit epoch wrapped: older words could match again
        hipError_t e = hipMemsetAsync(d_status, 0, sizeof(unsigned long long) * (size_t) capacity, st);
        if (e != hipSuccess) return e;
        epoch = 1;
    }
    c->status = d_status;
    c->tile_counter = d_counter;
    c->tile_base = 0;
    c->epoch = epoch;
    c->tiles = tiles;
    return hipMemsetAsync(d_counter, 0, sizeof(unsigned long long), st);
}

void ChainState::release()
{
    if (d_status) (void) hipFree(d_status);
    if (d_counter) (void) hipFree(d_counter);
    d_status = nullptr;
    d_counter = nullptr;
    capacity = 0;
}

}  // namespace b40c_hip

struct b40c_scan_enactor {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ChainState chain;
    float last_ms = 0.f;
};

namespace {

bfs_status timed_sync(b40c_scan_enactor *e)
{
    hipError_t r;
    if ((r = hipEventRecord(e->ev1, e->stream)) != hipSuccess) return host_hip_fail(r, "record");
    if ((r = hipEventSynchronize(e->ev1)) != hipSuccess) return host_hip_fail(r, "synchronize");
    if ((r = hipEventElapsedTime(&e->last_ms, e->ev0, e->ev1)) != hipSuccess) return host_hip_fail(r, "elapsed");
    return BFS_OK;
}

}  // namespace

extern "C" {

bfs_status b40c_scan_enactor_create(int32_t device, b40c_scan_enactor **out)
{
    if (!out) return host_fail(BFS_ERR_INVALID_ARG, "null output handle");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return host_fail(BFS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return host_fail(BFS_ERR_INVALID_ARG, "device ordinal out of range");
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return host_hip_fail(e, "hipSetDevice");
    b40c_scan_enactor *en = new b40c_scan_enactor();
    en->device = device;
    if ((e = hipStreamCreateWithFlags(&en->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&en->ev0)) != hipSuccess || (e = hipEventCreate(&en->ev1)) != hipSuccess ||
        (e = en->chain.init(device)) != hipSuccess) {
        b40c_scan_enactor_destroy(en);
        return host_hip_fail(e, "create scan enactor");
    }
    *out = en;
    return BFS_OK;
}

void b40c_scan_enactor_destroy(b40c_scan_enactor *e)
{
    if (!e) return;
    (void) hipSetDevice(e->device);
    if (e->stream) (void) hipStreamSynchronize(e->stream);
    e->chain.release();
    if (e->ev0) (void) hipEventDestroy(e->ev0);
    if (e->ev1) (void) hipEventDestroy(e->ev1);
    if (e->stream) (void) hipStreamDestroy(e->stream);
    delete e;
}

bfs_status b40c_scan(b40c_scan_enactor *e, const uint32_t *d_in, uint32_t *d_out, int64_t n, int32_t exclusive,
                     uint32_t *h_total)
{
    if (!e || n < 0 || (n > 0 && (!d_in || !d_out))) return host_fail(BFS_ERR_INVALID_ARG, "invalid scan arguments");
    if (n > ((int64_t) 1 << 40)) return host_fail(BFS_ERR_INVALID_ARG, "scan longer than 2^40 items");
    hipError_t r;
    if ((r = hipSetDevice(e->device)) != hipSuccess) return host_hip_fail(r, "hipSetDevice");
    if (n == 0) {
        if (h_total) *h_total = 0;
        e->last_ms = 0.f;
        return BFS_OK;
    }
    ChainArgs c;
    if ((r = e->chain.prepare(n, kScanTile, &c, e->stream)) != hipSuccess) return host_hip_fail(r, "scan status storage");
    if ((r = hipEventRecord(e->ev0, e->stream)) != hipSuccess) return host_hip_fail(r, "record");
    if ((r = launch_scan_u32(d_in, d_out, n, exclusive != 0, c, e->chain.d_word, e->stream)) != hipSuccess)
        return host_hip_fail(r, "launch scan");
    bfs_status st = timed_sync(e);
    if (st != BFS_OK) return st;
    if (h_total && (r = hipMemcpy(h_total, e->chain.d_word, sizeof(uint32_t), hipMemcpyDeviceToHost)) != hipSuccess)
        return host_hip_fail(r, "read scan total");
    return BFS_OK;
}

bfs_status b40c_copy_if(b40c_scan_enactor *e, const uint32_t *d_in, uint32_t *d_out, int64_t n, b40c_select_op op,
                        uint32_t comparand, int64_t *h_num_compacted, uint32_t *d_num_compacted)
{
    if (!e || n < 0 || (n > 0 && (!d_in || !d_out)) || op < B40C_SELECT_LESS || op > B40C_SELECT_NONNEGATIVE)
        return host_fail(BFS_ERR_INVALID_ARG, "invalid copy-if arguments");
    if (n > (int64_t) UINT32_MAX) return host_fail(BFS_ERR_INVALID_ARG, "copy-if longer than 2^32 - 1 items");
    hipError_t r;
    if ((r = hipSetDevice(e->device)) != hipSuccess) return host_hip_fail(r, "hipSetDevice");
    if (n == 0) {
        if (h_num_compacted) *h_num_compacted = 0;
        if (d_num_compacted && (r = hipMemsetAsync(d_num_compacted, 0, sizeof(uint32_t), e->stream)) != hipSuccess)
            return host_hip_fail(r, "clear count");
        if ((r = hipStreamSynchronize(e->stream)) != hipSuccess) return host_hip_fail(r, "synchronize");
        e->last_ms = 0.f;
        return BFS_OK;
    }
    ChainArgs c;
    if ((r = e->chain.prepare(n, kScanTile, &c, e->stream)) != hipSuccess) return host_hip_fail(r, "copy-if status storage");
    uint32_t *count = d_num_compacted ? d_num_compacted : e->chain.d_word;
    if ((r = hipEventRecord(e->ev0, e->stream)) != hipSuccess) return host_hip_fail(r, "record");
    if ((r = launch_copy_if_u32(d_in, d_out, n, (int) op, comparand, c, count, e->stream)) != hipSuccess)
        return host_hip_fail(r, "launch copy-if");
    bfs_status st = timed_sync(e);
    if (st != BFS_OK) return st;
    if (h_num_compacted) {
        uint32_t k = 0;
        if ((r = hipMemcpy(&k, count, sizeof(uint32_t), hipMemcpyDeviceToHost)) != hipSuccess)
            return host_hip_fail(r, "read compacted count");
        *h_num_compacted = k;
    }
    return BFS_OK;
}

bfs_status b40c_scan_noalloc(int32_t device, const uint32_t *d_in, uint32_t *d_out, int64_t n, int32_t exclusive,
                             uint32_t *d_total, void *stream)
{
    if (n < 0 || (n > 0 && (!d_in || !d_out))) return host_fail(BFS_ERR_INVALID_ARG, "invalid scan arguments");
    if (n > kNoAllocMaxTiles * kScanTile)
        return host_fail(BFS_ERR_INVALID_ARG, "scan longer than the allocation-free limit of 2^26 items");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return host_fail(BFS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count || device >= kNoAllocDevices)
        return host_fail(BFS_ERR_INVALID_ARG, "device ordinal out of range");
    hipError_t r;
    if ((r = hipSetDevice(device)) != hipSuccess) return host_hip_fail(r, "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n == 0) {
        if (d_total) {
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, d_total, 1, 0u);
            if ((r = hipGetLastError()) != hipSuccess) return host_hip_fail(r, "clear scan total");
        }
        return BFS_OK;
    }
    unsigned long long *words = nullptr;
    if ((r = noalloc_scratch(device, &words)) != hipSuccess) return host_hip_fail(r, "scan scratch address");
    ChainArgs c;
    c.tiles = (n + kScanTile - 1) / kScanTile;
    c.status = words + kNoAllocStatus;
    c.tile_counter = words;
    c.tile_base = 0;
    c.epoch = 0;
    const unsigned clear_blocks = (unsigned) std::min<long long>((c.tiles + kThreads - 1) / kThreads, 64);
    hipLaunchKernelGGL(noalloc_clear_kernel, dim3(clear_blocks), dim3(kThreads), 0, st, words, c.tiles);
    if ((r = hipGetLastError()) != hipSuccess) return host_hip_fail(r, "launch scratch clear");
    if ((r = launch_scan_u32(d_in, d_out, n, exclusive != 0, c, d_total, st)) != hipSuccess)
        return host_hip_fail(r, "launch scan");
    return BFS_OK;
}

bfs_status b40c_scan_last_ms(const b40c_scan_enactor *e, double *ms)
{
    if (!e || !ms) return host_fail(BFS_ERR_INVALID_ARG, "null argument");
    *ms = e->last_ms;
    return BFS_OK;
}

}  // extern "C"
