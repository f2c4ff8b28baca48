/*
 * b40c_scan.h -- single-pass prefix scan and stream compaction on the device
 * (C ABI, exported by libb40c_hip.so).
 *
 * These replace the reference's standalone scan and copy-if primitives,
 * which run three kernels per call (upsweep -> spine -> downsweep):
 *   b40c_scan_enactor          <- scan::Enactor (b40c/scan/enactor.cuh:54) and
 *                                 copy_if::Enactor (b40c/copy_if/enactor.cuh:52),
 *                                 which own their spine storage
 *   b40c_scan                  <- scan::Enactor::Scan (b40c/scan/enactor.cuh:466-561,
 *                                 pass at :426-454), Sum operator, EXCLUSIVE or not
 *                                 (b40c/scan/problem_type.cuh:49)
 *   b40c_copy_if               <- copy_if::Enactor::CopyIf (b40c/copy_if/enactor.cuh:
 *                                 485-571, pass at :436-461) with the select operator
 *                                 of test/copy-if/test_copy_if.h:50-60 (key < comparand)
 *                                 and two more fixed predicates
 * Here one kernel does each: every tile publishes its aggregate, then looks
 * back over its predecessors' published values (wave-parallel, 64 tiles per
 * probe) until it meets an inclusive prefix -- a decoupled look-back chained
 * scan.  Tiles take their index from an atomic counter in launch order, so a
 * tile only ever waits for tiles that are already running or done.
 * Copy-if is stable: selected items keep their input order.
 *
 * Calls are synchronous on the enactor's own stream.  The enactor owns its
 * tile-status scratch and grows it as needed; one enactor is used by one
 * host thread at a time.  Device pointers are plain device addresses.
 */
#ifndef B40C_SCAN_H
#define B40C_SCAN_H

#include <stdint.h>

#include "b40c_bfs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct b40c_scan_enactor b40c_scan_enactor;

typedef enum b40c_select_op {
    B40C_SELECT_LESS = 0,         /* keep key < comparand (test_copy_if.h:50-60) */
    B40C_SELECT_NOT_EQUAL = 1,    /* keep key != comparand */
    B40C_SELECT_NONNEGATIVE = 2   /* keep (int32_t) key >= 0, e.g. culled -1 slots of a queue */
} b40c_select_op;

bfs_status b40c_scan_enactor_create(int32_t device, b40c_scan_enactor **out);
void b40c_scan_enactor_destroy(b40c_scan_enactor *e);

/* d_out[i] = sum of d_in[0..i) (exclusive != 0) or d_in[0..i] (inclusive),
 * modulo 2^32.  d_out may equal d_in.  *h_total (optional) receives the sum
 * of all n items, modulo 2^32. */
bfs_status b40c_scan(b40c_scan_enactor *e, const uint32_t *d_in, uint32_t *d_out, int64_t n, int32_t exclusive,
                     uint32_t *h_total);

/* Copies the items of d_in[0..n) that satisfy the select operator to
 * d_out[0..count), in input order (d_out must not overlap d_in).  The count
 * goes to *h_num_compacted (host, optional) and *d_num_compacted (device,
 * optional). */
bfs_status b40c_copy_if(b40c_scan_enactor *e, const uint32_t *d_in, uint32_t *d_out, int64_t n,
                        b40c_select_op op, uint32_t comparand, int64_t *h_num_compacted,
                        uint32_t *d_num_compacted);

/* addition, no counterpart in the reference: b40c_scan that allocates
 * nothing, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The tile counter and status words
 * live in a static device array of 128 KiB per GPU, held until the process
 * ends; n may be at most 2^26 (67,108,864) items, larger calls return
 * BFS_ERR_INVALID_ARG.  Asynchronous: the clear and scan kernels are queued
 * on `stream` (a hipStream_t; null = the default stream) and the call
 * returns without waiting.  The sum of all n items goes to *d_total (device,
 * optional).  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (other streams or host threads) race and give
 * wrong results; calls queued on one stream are ordered and safe. */
bfs_status b40c_scan_noalloc(int32_t device, const uint32_t *d_in, uint32_t *d_out, int64_t n, int32_t exclusive,
                             uint32_t *d_total, void *stream);

/* addition: elapsed device time of the enactor's last call (ms, hipEvents). */
bfs_status b40c_scan_last_ms(const b40c_scan_enactor *e, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* B40C_SCAN_H */
