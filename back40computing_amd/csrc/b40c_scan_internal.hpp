// b40c_scan_internal.hpp -- host launchers of the single-pass scan,
// compaction and sort kernels, shared by the scan C ABI (b40c_scan.hip) and
// the device CSR builder (b40c_csr_build.hip).  Not installed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b40c_hip {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // items per tile of the chained kernels

// Arguments of one chained (decoupled look-back) launch.
struct ChainArgs {
    unsigned long long *status;        // one word per tile
    unsigned long long *tile_counter;  // launch-order tile indices, zeroed before the launch
    unsigned long long tile_base;      // counter value when this launch begins (0)
    unsigned epoch;                    // tags this launch's status words
    long long tiles;                   // tiles (= workgroups) of the launch
};

// Device storage of the chained launches of one owner (an enactor or a
// builder), used by one host thread and one stream at a time.
struct ChainState {
    int device = 0;
    unsigned long long *d_counter = nullptr;  // followed by 4 scratch words (d_word)
    uint32_t *d_word = nullptr;
    unsigned long long *d_status = nullptr;
    long long capacity = 0;                   // status words allocated
    unsigned epoch = 0;

    hipError_t init(int dev);
    // Status words and a fresh epoch for a launch over `items` in tiles of
    // tile_items, and the tile counter zeroed on `st` ahead of the launch.
    hipError_t prepare(long long items, int tile_items, ChainArgs *c, hipStream_t st);
    void release();
};

hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, long long n, bool exclusive, const ChainArgs &c,
                           uint32_t *d_total, hipStream_t st);
hipError_t launch_copy_if_u32(const uint32_t *in, uint32_t *out, long long n, int op, uint32_t comparand,
                              const ChainArgs &c, uint32_t *d_count, hipStream_t st);

}  // namespace b40c_hip
