// bfs_batch_internal.hpp -- host launchers of the batched kernels
// (bfs_batch_kernels.hip), used by the batch / server host code.
#pragma once

#include <hip/hip_runtime.h>

#include "bfs_batch.hpp"

namespace b40c_hip {

hipError_t launch_batch_persistent(const BatchArgs &a, int grid, hipStream_t st);
// One level (all queries): the level kernel, the grid-tier (hub) kernel and
// the single-workgroup advance kernel, kBatchLevelLaunches launches in all;
// *alive_out (device) receives the number of queries still running.
constexpr int kBatchLevelLaunches = 3;
hipError_t launch_batch_level(const BatchArgs &a, int grid, int *alive_out, hipStream_t st);
// Resets and seeds the queries whose slot indices are in d_list[0..count).
hipError_t launch_batch_seed(const BatchArgs &a, const int *d_list, int count, int max_nodes, hipStream_t st);
int batch_persistent_blocks_per_cu();
int batch_level_blocks_per_cu();

// Host helpers defined in bfs_capi.hip.
}  // namespace b40c_hip
#include "b40c_bfs.h"
namespace b40c_hip {
bfs_status host_fail(bfs_status s, const char *msg);
bfs_status host_hip_fail(hipError_t e, const char *what);
bfs_status host_validate_csr(const int32_t *row, const int32_t *col, int32_t nodes, int64_t edges);
bool host_csr_is_symmetric(const int32_t *row, const int32_t *col, int32_t nodes);

}  // namespace b40c_hip
