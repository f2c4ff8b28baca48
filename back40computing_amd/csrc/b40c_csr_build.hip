// b40c_csr_build.hip -- CSR construction on the device (C ABI in
// include/b40c_graph.h, b40c_device_csr_*).
//
// The reference builds every CSR on the host: CsrGraph::FromCoo stable-sorts
// the COO tuples by row (b40c/graph/csr_graph.cuh:108-152), after builders
// and file readers produced them (b40c/graph/builder/*.cuh).  Here the
// tuples become 64-bit keys (row << b | column, b = bits of V - 1) that an
// LSD radix sort orders on the device (8-bit digits, 2b bits in all: a
// per-tile digit histogram, one chained scan of the digit-major count matrix,
// then a stable scatter ranked by wave ballots); duplicates are dropped by a
// chained stream compaction; row offsets are written at the row boundaries
// of the sorted keys.  Rows come out sorted by column, as the host builder
// (graph_builder.cpp) makes them, so the two agree bit for bit.
//
// Self-loops to drop become the key of (2^b - 1, 2^b - 1), itself a loop, so
// they sort after every kept key and are cut off by count.
#include "b40c_graph.h"
#include "b40c_primitives.hpp"
#include "b40c_scan_internal.hpp"
#include "bfs_internal.hpp"
#include "graph_builder_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <vector>

using namespace b40c_hip;

namespace b40c_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / prim::kWave;
constexpr int kRounds = 16;                          // keys per lane
constexpr int kTile = kThreads * kRounds;            // 4096 keys per tile
constexpr int kWaveKeys = prim::kWave * kRounds;     // 1024 contiguous keys per wave
constexpr int kDigits = 256;
static_assert(kTile == kScanTile, "sort tiles match the chained-scan tiles");

__device__ __forceinline__ unsigned digit_of(unsigned long long key, int shift)
{
    return (unsigned) (key >> shift) & (kDigits - 1);
}

// Per-tile digit counts, digit-major: counts[d * tiles + tile].
__global__ __launch_bounds__(kThreads) void sort_upsweep_kernel(const unsigned long long *keys, long long n, int shift,
                                                                unsigned *counts, long long tiles)
{
    __shared__ unsigned hist[kDigits];
    hist[threadIdx.x] = 0u;
    __syncthreads();
    const long long tile = blockIdx.x;
    const int wave = threadIdx.x / prim::kWave, lane = prim::lane();
    const long long base = tile * kTile + (long long) wave * kWaveKeys + lane;
#pragma unroll 4
    for (int r = 0; r < kRounds; r++) {
        const long long i = base + (long long) r * prim::kWave;
        if (i < n) atomicAdd(&hist[digit_of(keys[i], shift)], 1u);
    }
    __syncthreads();
    counts[(long long) threadIdx.x * tiles + tile] = hist[threadIdx.x];
}

// Stable scatter by digit.  Wave w of a tile owns keys [w * 1024, (w + 1) *
// 1024) of the tile and walks them in 16 rounds of 64; a key's slot is its
// digit's scanned offset for this tile, plus the keys of that digit in the
// tile's lower waves, plus those earlier in its own wave (ballot match of
// the eight digit bits).
__global__ __launch_bounds__(kThreads) void sort_downsweep_kernel(const unsigned long long *keys_in,
                                                                  unsigned long long *keys_out, long long n, int shift,
                                                                  const unsigned *offsets, long long tiles)
{
    __shared__ unsigned base_of[kWaves][kDigits];
    unsigned *flat = &base_of[0][0];
    for (int i = threadIdx.x; i < kWaves * kDigits; i += kThreads) flat[i] = 0u;
    __syncthreads();
    const long long tile = blockIdx.x;
    const int wave = threadIdx.x / prim::kWave, lane = prim::lane();
    const long long base = tile * kTile + (long long) wave * kWaveKeys + lane;
    unsigned long long key[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const long long i = base + (long long) r * prim::kWave;
        key[r] = i < n ? keys_in[i] : 0ull;
        if (i < n) atomicAdd(&base_of[wave][digit_of(key[r], shift)], 1u);
    }
    __syncthreads();
    {
        // thread = digit: this tile's offset, then the lower waves' counts
        unsigned run = offsets[(long long) threadIdx.x * tiles + tile];
        for (int w = 0; w < kWaves; w++) {
            const unsigned c = base_of[w][threadIdx.x];
            base_of[w][threadIdx.x] = run;
            run += c;
        }
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const long long i = base + (long long) r * prim::kWave;
        const bool valid = i < n;
        const unsigned d = digit_of(key[r], shift);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        // every lane reads its digit's base before the group's first lane
        // advances it (one wave's LDS operations complete in order)
        const unsigned at = base_of[wave][d];
        const int rank = __popcll(peers & below);
        if (valid) {
            keys_out[at + (unsigned) rank] = key[r];
            if (rank == 0) base_of[wave][d] = at + (unsigned) __popcll(peers);
        }
    }
}

// Keys of (src, dst) pairs; with symmetrize also (dst, src); self-loops to
// drop become the loop key (2^b - 1, 2^b - 1) and are counted in *loops.
__global__ __launch_bounds__(kThreads) void pair_keys_kernel(const int *src, const int *dst, long long pairs, int bits,
                                                             int symmetrize, int drop_loops,
                                                             unsigned long long *keys, unsigned long long *loops)
{
    const unsigned long long loop_key = ((1ull << bits) - 1ull) << bits | ((1ull << bits) - 1ull);
    unsigned long long dropped = 0;
    for (long long i = blockIdx.x * (long long) kThreads + threadIdx.x; i < pairs;
         i += (long long) gridDim.x * kThreads) {
        const unsigned long long u = (unsigned) src[i], v = (unsigned) dst[i];
        const bool drop = drop_loops && u == v;
        const long long at = symmetrize ? 2 * i : i;
        keys[at] = drop ? loop_key : (u << bits) | v;
        if (symmetrize) keys[at + 1] = drop ? loop_key : (v << bits) | u;
        if (drop) dropped += symmetrize ? 2 : 1;
    }
    if (dropped) atomicAdd(loops, dropped);
}

// SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) and the counter-based stream of
// the host builders (graph_builder.cpp), so device samples equal host ones.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr unsigned long long kGolden = 0x9E3779B97F4A7C15ull;

// Graph500-style R-MAT samples (one per e), scrambled through perm, emitted
// as symmetric keys with loops dropped, exactly as b40c_graph_rmat draws them.
__global__ __launch_bounds__(kThreads) void rmat_keys_kernel(long long samples, int scale, unsigned long long seed,
                                                             unsigned long long t_a, unsigned long long t_ab,
                                                             unsigned long long t_abc, const int *perm, int bits,
                                                             unsigned long long *keys, unsigned long long *loops)
{
    const unsigned long long loop_key = ((1ull << bits) - 1ull) << bits | ((1ull << bits) - 1ull);
    unsigned long long dropped = 0;
    for (long long e = blockIdx.x * (long long) kThreads + threadIdx.x; e < samples;
         e += (long long) gridDim.x * kThreads) {
        unsigned long long s = mix64(seed ^ mix64((unsigned long long) e * kGolden + 0x632BE59BD9B4E019ull));
        unsigned u = 0, v = 0;
        unsigned long long draw = 0;
        for (int lvl = 0; lvl < scale; lvl++) {
            if ((lvl & 1) == 0) {
                s += kGolden;
                draw = mix64(s);
            }
            const unsigned long long r = (lvl & 1) ? (draw >> 32) : (draw & 0xFFFFFFFFull);
            unsigned ub, vb;
            if (r < t_a) { ub = 0; vb = 0; }
            else if (r < t_ab) { ub = 0; vb = 1; }
            else if (r < t_abc) { ub = 1; vb = 0; }
            else { ub = 1; vb = 1; }
            u = (u << 1) | ub;
            v = (v << 1) | vb;
        }
        if (perm) {
            u = (unsigned) perm[u];
            v = (unsigned) perm[v];
        }
        const bool drop = u == v;
        keys[2 * e] = drop ? loop_key : ((unsigned long long) u << bits) | v;
        keys[2 * e + 1] = drop ? loop_key : ((unsigned long long) v << bits) | u;
        if (drop) dropped += 2;
    }
    if (dropped) atomicAdd(loops, dropped);
}

// Chained stable compaction of sorted keys: keeps key i when it differs from
// key i - 1 (consecutive duplicates collapse to the first).
__global__ __launch_bounds__(kThreads) void unique_kernel(const unsigned long long *in, unsigned long long *out,
                                                          long long n, ChainArgs c, unsigned *d_count)
{
    __shared__ long long s_tile;
    __shared__ unsigned s_wave[kWaves], s_prefix;
    if (threadIdx.x == 0) s_tile = (long long) (atomicAdd(c.tile_counter, 1ull) - c.tile_base);
    __syncthreads();
    const long long tile = s_tile;
    const long long base = tile * kTile + (long long) threadIdx.x * kRounds;
    unsigned long long v[kRounds];
    unsigned long long prev = base > 0 && base - 1 < n ? in[base - 1] : 0ull;
    unsigned keep = 0;
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
        const long long i = base + k;
        v[k] = i < n ? in[i] : 0ull;
        if (i < n && (i == 0 || v[k] != prev)) keep |= 1u << k;
        prev = v[k];
    }
    unsigned agg;
    const unsigned before = prim::block_exclusive_sum_u32<kThreads>((unsigned) __popc(keep), s_wave, &agg);
    if (threadIdx.x < prim::kWave) {
        unsigned prefix = 0;
        if (tile == 0) {
            if (threadIdx.x == 0) prim::publish(c.status, tile, c.epoch, prim::kStatePrefix, agg);
        } else {
            if (threadIdx.x == 0) prim::publish(c.status, tile, c.epoch, prim::kStateAggregate, agg);
            prefix = prim::lookback_prefix(c.status, tile, c.epoch);
            if (threadIdx.x == 0) prim::publish(c.status, tile, c.epoch, prim::kStatePrefix, prefix + agg);
        }
        if (threadIdx.x == 0) s_prefix = prefix;
    }
    __syncthreads();
    unsigned at = s_prefix + before;
#pragma unroll
    for (int k = 0; k < kRounds; k++)
        if ((keep >> k) & 1u) out[at++] = v[k];
    if (tile == c.tiles - 1 && threadIdx.x == 0) *d_count = s_prefix + agg;
}

// Columns and row offsets from E sorted keys: key i starts row u = key >> b,
// and writes row[w] = i for every w in (row of key i - 1, u]; the last key
// closes the rows after it with E.
__global__ __launch_bounds__(kThreads) void csr_from_keys_kernel(const unsigned long long *keys, long long edges,
                                                                 int bits, int nodes, int *row, int *col)
{
    const unsigned long long mask = (1ull << bits) - 1ull;
    for (long long i = blockIdx.x * (long long) kThreads + threadIdx.x; i < edges;
         i += (long long) gridDim.x * kThreads) {
        const unsigned long long k = keys[i];
        const int u = (int) (k >> bits);
        col[i] = (int) (k & mask);
        const int prev = i == 0 ? -1 : (int) (keys[i - 1] >> bits);
        for (int w = prev + 1; w <= u; w++) row[w] = (int) i;
        if (i == edges - 1)
            for (int w = u + 1; w <= nodes; w++) row[w] = (int) edges;
    }
}

__global__ __launch_bounds__(kThreads) void fill_int_kernel(int *p, long long n, int value)
{
    for (long long i = blockIdx.x * (long long) kThreads + threadIdx.x; i < n; i += (long long) gridDim.x * kThreads)
        p[i] = value;
}

// (column, row) keys of every edge of a CSR: the transpose's tuples.  Each
// edge finds its row by a binary search of the row offsets.
__global__ __launch_bounds__(kThreads) void transpose_keys_kernel(const int *row, const int *col, int nodes,
                                                                  long long edges, int bits,
                                                                  unsigned long long *keys)
{
    for (long long e = blockIdx.x * (long long) kThreads + threadIdx.x; e < edges;
         e += (long long) gridDim.x * kThreads) {
        int lo = 0, hi = nodes - 1;  // last u with row[u] <= e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((long long) row[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        keys[e] = ((unsigned long long) (unsigned) col[e] << bits) | (unsigned) lo;
    }
}

unsigned grid_for(long long n)
{
    return (unsigned) std::max(1LL, std::min((n + kThreads - 1) / kThreads, 16384LL));
}

int bits_for(int nodes)
{
    int b = 1;
    while (b < 31 && (1LL << b) < (long long) nodes) b++;
    return b;
}

// Device storage and stream of one build.
struct Builder {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ChainState chain;
    unsigned long long *keys[2] = {nullptr, nullptr};
    unsigned *counts = nullptr;
    int *perm = nullptr;
    int *src = nullptr, *dst = nullptr;
    int status = 0;  // first failure: 1 invalid, 2 out of memory, 4 device error

    ~Builder()
    {
        if (stream) (void) hipStreamSynchronize(stream);
        for (auto *k : keys)
            if (k) (void) hipFree(k);
        if (counts) (void) hipFree(counts);
        if (perm) (void) hipFree(perm);
        if (src) (void) hipFree(src);
        if (dst) (void) hipFree(dst);
        chain.release();
        if (ev0) (void) hipEventDestroy(ev0);
        if (ev1) (void) hipEventDestroy(ev1);
        if (stream) (void) hipStreamDestroy(stream);
    }

    bool ok(hipError_t e)
    {
        if (e == hipSuccess) return true;
        if (!status) status = e == hipErrorOutOfMemory ? 2 : 4;
        return false;
    }

    bool start(int dev)
    {
        device = dev;
        return ok(hipSetDevice(dev)) && ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) &&
               ok(hipEventCreate(&ev0)) && ok(hipEventCreate(&ev1)) && ok(chain.init(dev)) &&
               ok(hipEventRecord(ev0, stream));
    }

    bool alloc_keys(long long n)
    {
        const size_t b = sizeof(unsigned long long) * (size_t) std::max(1LL, n);
        return ok(hipMalloc(&keys[0], b)) && ok(hipMalloc(&keys[1], b));
    }

    // LSD radix sort of keys[0][0, n) over bits [0, key_bits); the result
    // is in keys[0].
    bool sort(long long n, int key_bits)
    {
        if (n <= 1) return true;
        const long long tiles = (n + kTile - 1) / kTile;
        if (!ok(hipMalloc(&counts, sizeof(unsigned) * (size_t) (tiles * kDigits)))) return false;
        int cur = 0;
        for (int shift = 0; shift < key_bits; shift += 8) {
            hipLaunchKernelGGL(sort_upsweep_kernel, dim3((unsigned) tiles), dim3(kThreads), 0, stream, keys[cur], n,
                               shift, counts, tiles);
            if (!ok(hipGetLastError())) return false;
            ChainArgs c;
            if (!ok(chain.prepare(tiles * kDigits, kScanTile, &c, stream))) return false;
            if (!ok(launch_scan_u32(counts, counts, tiles * kDigits, true, c, nullptr, stream))) return false;
            hipLaunchKernelGGL(sort_downsweep_kernel, dim3((unsigned) tiles), dim3(kThreads), 0, stream, keys[cur],
                               keys[cur ^ 1], n, shift, counts, tiles);
            if (!ok(hipGetLastError())) return false;
            cur ^= 1;
        }
        if (cur == 1) std::swap(keys[0], keys[1]);
        (void) hipFree(counts);
        counts = nullptr;
        return true;
    }

    // Number of kept keys after dropping `loops` trailing loop keys and, with
    // dedup, consecutive duplicates; the kept keys end up in keys[0].
    bool compact(long long n, long long loops, bool dedup, long long *kept)
    {
        long long m = n - loops;
        if (dedup && m > 0) {
            ChainArgs c;
            if (!ok(chain.prepare(m, kScanTile, &c, stream))) return false;
            hipLaunchKernelGGL(unique_kernel, dim3((unsigned) c.tiles), dim3(kThreads), 0, stream, keys[0], keys[1],
                               m, c, chain.d_word);
            if (!ok(hipGetLastError())) return false;
            unsigned k = 0;
            if (!ok(hipMemcpyAsync(&k, chain.d_word, sizeof(unsigned), hipMemcpyDeviceToHost, stream)) ||
                !ok(hipStreamSynchronize(stream)))
                return false;
            m = k;
            std::swap(keys[0], keys[1]);
        }
        *kept = m;
        return true;
    }

    bool finish(int nodes, long long edges, int bits, bool symmetric, b40c_device_csr *out)
    {
        if (edges > INT32_MAX) {
            if (!status) status = 3;
            return false;
        }
        int *row = nullptr, *col = nullptr;
        if (!ok(hipMalloc(&row, sizeof(int) * ((size_t) nodes + 1))) ||
            !ok(hipMalloc(&col, sizeof(int) * (size_t) std::max(1LL, edges)))) {
            if (row) (void) hipFree(row);
            return false;
        }
        if (edges == 0) hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(nodes + 1)), dim3(kThreads), 0, stream, row,
                                           (long long) nodes + 1, 0);
        else hipLaunchKernelGGL(csr_from_keys_kernel, dim3(grid_for(edges)), dim3(kThreads), 0, stream, keys[0],
                                edges, bits, nodes, row, col);
        float ms = 0.f;
        if (!ok(hipGetLastError()) || !ok(hipEventRecord(ev1, stream)) || !ok(hipEventSynchronize(ev1)) ||
            !ok(hipEventElapsedTime(&ms, ev0, ev1))) {
            (void) hipFree(row);
            (void) hipFree(col);
            return false;
        }
        out->device = device;
        out->nodes = nodes;
        out->edges = edges;
        out->row_offsets = row;
        out->column_indices = col;
        out->symmetric = symmetric ? 1 : 0;
        out->build_ms = ms;
        return true;
    }
};

int device_ok(int32_t device)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return 4;
    return device >= 0 && device < count ? 0 : 1;
}

}  // namespace
}  // namespace b40c_hip

extern "C" {

int b40c_device_csr_from_coo(int32_t device, int32_t nodes, int64_t num_pairs, const int32_t *h_src,
                             const int32_t *h_dst, int symmetrize, int dedup, int drop_self_loops,
                             b40c_device_csr *out)
{
    if (!out || nodes <= 0 || num_pairs < 0 || (num_pairs > 0 && (!h_src || !h_dst))) return 1;
    *out = b40c_device_csr{};
    for (int64_t i = 0; i < num_pairs; i++)
        if (h_src[i] < 0 || h_src[i] >= nodes || h_dst[i] < 0 || h_dst[i] >= nodes) return 1;
    const long long n = (symmetrize ? 2 : 1) * (long long) num_pairs;
    if (n > (long long) UINT32_MAX) return 3;  // sort offsets are 32-bit
    if (int rc = device_ok(device)) return rc;
    Builder b;
    if (!b.start(device)) return b.status;
    const int bits = bits_for(nodes);
    long long edges = 0;
    if (n > 0) {
        unsigned long long *d_loops = reinterpret_cast<unsigned long long *>(b.chain.d_word);  // 8-byte aligned
        unsigned long long loops = 0;
        const size_t pb = sizeof(int) * (size_t) num_pairs;
        if (!b.alloc_keys(n) || !b.ok(hipMalloc(&b.src, pb)) || !b.ok(hipMalloc(&b.dst, pb)) ||
            !b.ok(hipMemcpyAsync(b.src, h_src, pb, hipMemcpyHostToDevice, b.stream)) ||
            !b.ok(hipMemcpyAsync(b.dst, h_dst, pb, hipMemcpyHostToDevice, b.stream)) ||
            !b.ok(hipMemsetAsync(d_loops, 0, sizeof(unsigned long long), b.stream)))
            return b.status;
        hipLaunchKernelGGL(pair_keys_kernel, dim3(grid_for(num_pairs)), dim3(kThreads), 0, b.stream, b.src, b.dst,
                           (long long) num_pairs, bits, symmetrize ? 1 : 0, drop_self_loops ? 1 : 0, b.keys[0],
                           d_loops);
        if (!b.ok(hipGetLastError()) ||
            !b.ok(hipMemcpyAsync(&loops, d_loops, sizeof(loops), hipMemcpyDeviceToHost, b.stream)) ||
            !b.ok(hipStreamSynchronize(b.stream)))
            return b.status;
        if (!b.sort(n, 2 * bits) || !b.compact(n, (long long) loops, dedup != 0, &edges)) return b.status;
    }
    if (!b.finish(nodes, edges, bits, symmetrize != 0, out)) return b.status ? b.status : 4;
    return 0;
}

int b40c_device_csr_rmat(int32_t device, int32_t scale, int32_t edge_factor, double a, double b, double c,
                         uint64_t seed, int scramble, b40c_device_csr *out)
{
    if (!out || scale < 1 || scale > 30 || edge_factor < 1) return 1;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return 1;
    *out = b40c_device_csr{};
    const int32_t nodes = (int32_t) (1u << scale);
    const long long samples = (long long) edge_factor << scale;
    if (2 * samples > (long long) INT32_MAX * 2) return 3;
    if (int rc = device_ok(device)) return rc;
    const uint64_t two32 = 1ull << 32;
    const uint64_t t_a = (uint64_t) (a * two32), t_ab = (uint64_t) ((a + b) * two32),
                   t_abc = (uint64_t) ((a + b + c) * two32);
    std::vector<int32_t> perm;
    if (scramble) b40c_rmat_permutation(scale, seed, &perm);
    Builder bld;
    if (!bld.start(device)) return bld.status;
    const int bits = scale;
    const long long n = 2 * samples;
    unsigned long long *d_loops = reinterpret_cast<unsigned long long *>(bld.chain.d_word);
    unsigned long long loops = 0;
    if (!bld.alloc_keys(n) || !bld.ok(hipMemsetAsync(d_loops, 0, sizeof(unsigned long long), bld.stream)))
        return bld.status;
    if (scramble && (!bld.ok(hipMalloc(&bld.perm, sizeof(int) * (size_t) nodes)) ||
                     !bld.ok(hipMemcpyAsync(bld.perm, perm.data(), sizeof(int) * (size_t) nodes,
                                            hipMemcpyHostToDevice, bld.stream))))
        return bld.status;
    hipLaunchKernelGGL(rmat_keys_kernel, dim3(grid_for(samples)), dim3(kThreads), 0, bld.stream, samples, scale,
                       (unsigned long long) seed, (unsigned long long) t_a, (unsigned long long) t_ab,
                       (unsigned long long) t_abc, bld.perm, bits, bld.keys[0], d_loops);
    if (!bld.ok(hipGetLastError()) ||
        !bld.ok(hipMemcpyAsync(&loops, d_loops, sizeof(loops), hipMemcpyDeviceToHost, bld.stream)) ||
        !bld.ok(hipStreamSynchronize(bld.stream)))
        return bld.status;
    long long edges = 0;
    if (!bld.sort(n, 2 * bits) || !bld.compact(n, (long long) loops, true, &edges)) return bld.status;
    if (!bld.finish(nodes, edges, bits, true, out)) return bld.status ? bld.status : 4;
    return 0;
}

int b40c_device_csr_transpose(const b40c_device_csr *g, b40c_device_csr *out)
{
    if (!g || !out || g->nodes <= 0 || g->edges < 0 || !g->row_offsets) return 1;
    *out = b40c_device_csr{};
    if (int rc = device_ok(g->device)) return rc;
    Builder b;
    if (!b.start(g->device)) return b.status;
    const int bits = bits_for(g->nodes);
    const long long n = g->edges;
    if (n > 0) {
        if (!b.alloc_keys(n)) return b.status;
        hipLaunchKernelGGL(transpose_keys_kernel, dim3(grid_for(n)), dim3(kThreads), 0, b.stream, g->row_offsets,
                           g->column_indices, g->nodes, n, bits, b.keys[0]);
        if (!b.ok(hipGetLastError()) || !b.sort(n, 2 * bits)) return b.status;
    }
    if (!b.finish(g->nodes, n, bits, g->symmetric != 0, out)) return b.status ? b.status : 4;
    return 0;
}

int b40c_device_csr_to_host(const b40c_device_csr *g, b40c_csr *out)
{
    if (!g || !out || g->nodes <= 0 || g->edges < 0 || !g->row_offsets) return 1;
    if (hipSetDevice(g->device) != hipSuccess) return 4;
    out->nodes = g->nodes;
    out->edges = g->edges;
    out->row_offsets = (int32_t *) std::malloc(sizeof(int32_t) * ((size_t) g->nodes + 1));
    out->column_indices = (int32_t *) std::malloc(sizeof(int32_t) * (size_t) std::max<int64_t>(g->edges, 1));
    if (!out->row_offsets || !out->column_indices) {
        b40c_graph_free(out);
        return 2;
    }
    if (hipMemcpy(out->row_offsets, g->row_offsets, sizeof(int32_t) * ((size_t) g->nodes + 1),
                  hipMemcpyDeviceToHost) != hipSuccess ||
        (g->edges > 0 && hipMemcpy(out->column_indices, g->column_indices, sizeof(int32_t) * (size_t) g->edges,
                                   hipMemcpyDeviceToHost) != hipSuccess)) {
        b40c_graph_free(out);
        return 4;
    }
    return 0;
}

void b40c_device_csr_free(b40c_device_csr *g)
{
    if (!g) return;
    if (g->row_offsets || g->column_indices) (void) hipSetDevice(g->device);
    if (g->row_offsets) (void) hipFree(g->row_offsets);
    if (g->column_indices) (void) hipFree(g->column_indices);
    g->row_offsets = g->column_indices = nullptr;
    g->nodes = 0;
    g->edges = 0;
}

}  // extern "C"
