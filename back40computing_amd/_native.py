"""ctypes binding of libb40c_hip.so (C ABI in include/b40c_bfs.h, include/b40c_graph.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C back40computing_amd/csrc``).  Loading fails loudly when it is
missing: there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libb40c_hip.so")

c_int32_p = ctypes.POINTER(ctypes.c_int32)


class CsrStruct(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int32),
        ("edges", ctypes.c_int64),
        ("row_offsets", c_int32_p),
        ("column_indices", c_int32_p),
    ]


class CooStruct(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int32),
        ("pairs", ctypes.c_int64),
        ("src", c_int32_p),
        ("dst", c_int32_p),
    ]


class DeviceCsrStruct(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("nodes", ctypes.c_int32),
        ("edges", ctypes.c_int64),
        ("row_offsets", ctypes.c_void_p),
        ("column_indices", ctypes.c_void_p),
        ("symmetric", ctypes.c_int32),
        ("build_ms", ctypes.c_double),
    ]


# BFS_FAST_VISITED_MAX (include/b40c_bfs.h): most vertices for which the fast
# path keeps its copy of the visited bitmap in LDS (config fast_visited = 1).
FAST_VISITED_MAX = 1103872


class BfsConfig(ctypes.Structure):
    _fields_ = [
        ("wave_gather_threshold", ctypes.c_int32),
        ("cta_gather_threshold", ctypes.c_int32),
        ("fast_path_threshold", ctypes.c_int32),
        ("saturation_threshold", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("cull_mode", ctypes.c_int32),
        ("instrument", ctypes.c_int32),
        ("grid_gather_threshold", ctypes.c_int32),
        ("do_alpha", ctypes.c_int32),
        ("do_beta", ctypes.c_int32),
        ("tile_grab_max", ctypes.c_int32),
        ("bu_tile_blocks", ctypes.c_int32),
        ("grab_strided", ctypes.c_int32),
        ("claim_filter", ctypes.c_int32),
        ("spread_steps", ctypes.c_int32),
        ("compaction", ctypes.c_int32),
        ("fast_visited", ctypes.c_int32),
    ]


class BfsStats(ctypes.Structure):
    _fields_ = [
        ("total_queued", ctypes.c_int64),
        ("search_depth", ctypes.c_int32),
        ("avg_duty", ctypes.c_double),
        ("elapsed_ms", ctypes.c_double),
        ("kernel_launches", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("elapsed_ms", ctypes.c_double),
        ("queries", ctypes.c_int32),
        ("levels", ctypes.c_int32),
        ("launches", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("latency_p50_ms", ctypes.c_double),
        ("latency_p99_ms", ctypes.c_double),
        ("latency_max_ms", ctypes.c_double),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("max_slots", ctypes.c_int32),
        ("levels_per_tick", ctypes.c_int32),
        ("hbm_budget_bytes", ctypes.c_int64),
        ("direction_optimizing", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 5),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("ticket", ctypes.c_int64),
        ("graph_id", ctypes.c_int32),
        ("src", ctypes.c_int32),
        ("levels", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("submit_ms", ctypes.c_double),
        ("admit_ms", ctypes.c_double),
        ("finish_ms", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_int64),
        ("admitted", ctypes.c_int64),
        ("completed", ctypes.c_int64),
        ("rejected", ctypes.c_int64),
        ("ticks", ctypes.c_int64),
        ("max_resident", ctypes.c_int64),
        ("busy_ms", ctypes.c_double),
        ("tick_ms", ctypes.c_double),
        ("copy_ms", ctypes.c_double),
    ]


_lib = None


def _sig(lib, name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)


def load():
    """Returns the loaded library, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libb40c_hip.so is not built (expected at %s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    i64 = ctypes.c_int64
    u64 = ctypes.c_uint64
    dbl = ctypes.c_double
    csr_p = ctypes.POINTER(CsrStruct)
    _sig(lib, "b40c_graph_grid2d", ctypes.c_int, i32, csr_p)
    _sig(lib, "b40c_graph_grid3d", ctypes.c_int, i32, csr_p)
    _sig(lib, "b40c_graph_rmat", ctypes.c_int, i32, i32, dbl, dbl, dbl, u64, ctypes.c_int, csr_p)
    _sig(lib, "b40c_graph_erdos_renyi", ctypes.c_int, i32, i64, u64, csr_p)
    _sig(lib, "b40c_graph_random", ctypes.c_int, i32, i64, u64, ctypes.c_int, csr_p)
    _sig(lib, "b40c_graph_from_coo", ctypes.c_int, i32, i64, c_int32_p, c_int32_p,
         ctypes.c_int, ctypes.c_int, ctypes.c_int, csr_p)
    _sig(lib, "b40c_graph_sample_sources", ctypes.c_int, csr_p, i32, u64, c_int32_p)
    _sig(lib, "b40c_graph_free", None, csr_p)
    coo_p = ctypes.POINTER(CooStruct)
    _sig(lib, "b40c_read_dimacs", ctypes.c_int, ctypes.c_char_p, ctypes.c_int, coo_p)
    _sig(lib, "b40c_read_metis", ctypes.c_int, ctypes.c_char_p, coo_p)
    _sig(lib, "b40c_read_market", ctypes.c_int, ctypes.c_char_p, ctypes.c_int, coo_p)
    _sig(lib, "b40c_coo_free", None, coo_p)
    dcsr_p = ctypes.POINTER(DeviceCsrStruct)
    _sig(lib, "b40c_device_csr_from_coo", ctypes.c_int, i32, i32, i64, c_int32_p, c_int32_p,
         ctypes.c_int, ctypes.c_int, ctypes.c_int, dcsr_p)
    _sig(lib, "b40c_device_csr_rmat", ctypes.c_int, i32, i32, i32, dbl, dbl, dbl, u64, ctypes.c_int, dcsr_p)
    _sig(lib, "b40c_device_csr_transpose", ctypes.c_int, dcsr_p, dcsr_p)
    _sig(lib, "b40c_device_csr_to_host", ctypes.c_int, dcsr_p, csr_p)
    _sig(lib, "b40c_device_csr_free", None, dcsr_p)
    _sig(lib, "b40c_scan_enactor_create", ctypes.c_int, i32, ctypes.POINTER(vp))
    _sig(lib, "b40c_scan_enactor_destroy", None, vp)
    _sig(lib, "b40c_scan", ctypes.c_int, vp, vp, vp, i64, i32, ctypes.POINTER(ctypes.c_uint32))
    _sig(lib, "b40c_copy_if", ctypes.c_int, vp, vp, vp, i64, ctypes.c_int, ctypes.c_uint32,
         ctypes.POINTER(i64), vp)
    _sig(lib, "b40c_scan_last_ms", ctypes.c_int, vp, ctypes.POINTER(dbl))
    _sig(lib, "b40c_scan_noalloc", ctypes.c_int, i32, vp, vp, i64, i32, vp, vp)

    _sig(lib, "bfs_status_string", ctypes.c_char_p, ctypes.c_int)
    _sig(lib, "bfs_last_error_message", ctypes.c_char_p)
    _sig(lib, "bfs_strategy_frontier_type", ctypes.c_int, ctypes.c_int)
    _sig(lib, "bfs_config_default", ctypes.c_int, ctypes.POINTER(BfsConfig))
    _sig(lib, "bfs_device_count", ctypes.c_int, c_int32_p)
    _sig(lib, "bfs_device_synchronize", ctypes.c_int, i32)
    _sig(lib, "bfs_problem_create", ctypes.c_int, c_int32_p, c_int32_p, i32, i64, i32, i32,
         ctypes.POINTER(vp))
    _sig(lib, "bfs_problem_reset", ctypes.c_int, vp, ctypes.c_int, dbl)
    _sig(lib, "bfs_problem_set_config", ctypes.c_int, vp, ctypes.POINTER(BfsConfig))
    _sig(lib, "bfs_problem_get_config", ctypes.c_int, vp, ctypes.POINTER(BfsConfig))
    _sig(lib, "bfs_enact", ctypes.c_int, vp, ctypes.c_int, i32, i32, i32, ctypes.POINTER(BfsStats))
    _sig(lib, "bfs_extract_labels", ctypes.c_int, vp, c_int32_p)
    _sig(lib, "bfs_extract_depths", ctypes.c_int, vp, c_int32_p)
    _sig(lib, "bfs_problem_destroy", None, vp)
    _sig(lib, "bfs_problem_device_info", ctypes.c_int, vp, c_int32_p, c_int32_p,
         ctypes.POINTER(ctypes.c_int64))
    _sig(lib, "bfs_problem_device_arrays", ctypes.c_int, vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
         ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp))
    _sig(lib, "bfs_get_level_log", ctypes.c_int, vp, i32, ctypes.POINTER(ctypes.c_uint32),
         ctypes.POINTER(dbl), c_int32_p)
    _sig(lib, "bfs_barrier_benchmark", ctypes.c_int, i32, i32, i32, ctypes.POINTER(dbl), c_int32_p)
    _sig(lib, "bfs_copy_calibration", ctypes.c_int, i32, i64, ctypes.POINTER(dbl))
    vpp = ctypes.POINTER(vp)
    _sig(lib, "bfs_pool_create", ctypes.c_int, i32, vpp)
    _sig(lib, "bfs_pool_add_graph", ctypes.c_int, vp, c_int32_p, c_int32_p, i32, i64, c_int32_p)
    _sig(lib, "bfs_pool_add_device_graph", ctypes.c_int, vp, dcsr_p, c_int32_p)
    _sig(lib, "bfs_pool_graph_info", ctypes.c_int, vp, i32, c_int32_p, ctypes.POINTER(i64))
    _sig(lib, "bfs_pool_device_bytes", ctypes.c_int, vp, ctypes.POINTER(i64))
    _sig(lib, "bfs_pool_destroy", None, vp)
    _sig(lib, "bfs_query_state_bytes", i64, i32)
    _sig(lib, "bfs_batch_create", ctypes.c_int, vp, i32, vpp)
    _sig(lib, "bfs_batch_set_config", ctypes.c_int, vp, ctypes.POINTER(BfsConfig))
    _sig(lib, "bfs_batch_config_default", ctypes.c_int, ctypes.POINTER(BfsConfig))
    _sig(lib, "bfs_batch_enact", ctypes.c_int, vp, c_int32_p, c_int32_p, i32, i32, i32, ctypes.POINTER(BatchStats))
    _sig(lib, "bfs_batch_query_latency", ctypes.c_int, vp, ctypes.POINTER(dbl))
    _sig(lib, "bfs_batch_extract_depths", ctypes.c_int, vp, i32, c_int32_p)
    _sig(lib, "bfs_batch_destroy", None, vp)
    _sig(lib, "bfs_batch_level_log", ctypes.c_int, vp, i32, ctypes.POINTER(ctypes.c_uint32),
         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(dbl), c_int32_p)
    _sig(lib, "bfs_server_create", ctypes.c_int, vp, ctypes.POINTER(ServerConfig), vpp)
    _sig(lib, "bfs_server_submit", ctypes.c_int, vp, i32, i32, c_int32_p, ctypes.POINTER(i64))
    _sig(lib, "bfs_server_wait", ctypes.c_int, vp, i64, i32, ctypes.POINTER(RequestInfo))
    _sig(lib, "bfs_server_get_stats", ctypes.c_int, vp, ctypes.POINTER(ServerStats))
    _sig(lib, "bfs_server_destroy", None, vp)
    _sig(lib, "bfs_host_alloc_pinned", ctypes.c_int, i64, ctypes.POINTER(ctypes.c_void_p))
    _sig(lib, "bfs_host_free_pinned", None, vp)
    _sig(lib, "bfs_selftest_wave_primitives", ctypes.c_int, i32, c_int32_p)
    _lib = lib
    return lib


def i32ptr(a):
    """Pointer to the data of a C-contiguous int32 numpy array."""
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_int32_p)
