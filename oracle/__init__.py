"""CPU oracle for the BFS path -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle_bfs.so (built from oracle/bfs_oracle.c by
``make -C oracle``).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use it; results are compared against it, and the
product never calls it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_bfs.so")
_lib = None
_i32p = ctypes.POINTER(ctypes.c_int32)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle library missing (%s); run make -C oracle" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_bfs.restype = ctypes.c_int32
    lib.oracle_bfs.argtypes = [ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, _i32p, _i32p,
                               ctypes.POINTER(ctypes.c_double)]
    lib.oracle_edges_visited.restype = ctypes.c_int64
    lib.oracle_edges_visited.argtypes = [ctypes.c_int32, _i32p, _i32p, ctypes.POINTER(ctypes.c_int64)]
    lib.oracle_compare.restype = ctypes.c_int64
    lib.oracle_compare.argtypes = [_i32p, _i32p, ctypes.c_int64]
    lib.oracle_check_preds.restype = ctypes.c_int64
    lib.oracle_check_preds.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, _i32p]
    lib.oracle_level_counts.restype = ctypes.c_int32
    lib.oracle_level_counts.argtypes = [ctypes.c_int32, _i32p, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    _lib = lib
    return lib


def _p(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_i32p)


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def bfs(row_offsets, column_indices, src, with_preds=False):
    """Serial reference BFS.  Returns (labels, search_depth, elapsed_ms[, preds])."""
    row, col = _arr(row_offsets), _arr(column_indices)
    n = len(row) - 1
    labels = np.empty(n, dtype=np.int32)
    preds = np.empty(n, dtype=np.int32) if with_preds else None
    ms = ctypes.c_double()
    colp = _p(col) if len(col) else _p(np.zeros(1, np.int32))
    depth = load().oracle_bfs(n, _p(row), colp, int(src), _p(labels),
                              _p(preds) if with_preds else None, ctypes.byref(ms))
    if depth < 0:
        raise ValueError("oracle_bfs: invalid arguments")
    if with_preds:
        return labels, depth, ms.value, preds
    return labels, depth, ms.value


def edges_visited(row_offsets, labels):
    """(edges_visited, nodes_visited) as the reference's TEPS accounting."""
    nv = ctypes.c_int64()
    row, lab = _arr(row_offsets), _arr(labels)
    ev = load().oracle_edges_visited(len(row) - 1, _p(row), _p(lab), ctypes.byref(nv))
    return int(ev), int(nv.value)


def compare(computed, reference):
    """-1 if equal, else the first mismatching index."""
    a, b = _arr(computed), _arr(reference)
    assert len(a) == len(b)
    return int(load().oracle_compare(_p(a), _p(b), len(a)))


def check_preds(row_offsets, column_indices, depth, preds):
    """-1 if every predecessor is valid, else the first failing vertex."""
    row, col, d, p = _arr(row_offsets), _arr(column_indices), _arr(depth), _arr(preds)
    colp = _p(col) if len(col) else _p(np.zeros(1, np.int32))
    return int(load().oracle_check_preds(len(row) - 1, _p(row), colp, _p(d), _p(p)))


def level_counts(labels, max_depth=None):
    lab = _arr(labels)
    if max_depth is None:
        max_depth = int(lab.max()) + 1 if len(lab) and lab.max() >= 0 else 1
    counts = np.zeros(max_depth, dtype=np.int64)
    levels = load().oracle_level_counts(len(lab), _p(lab), counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        max_depth)
    if levels < 0:
        raise ValueError("label exceeds max_depth")
    return counts[:levels]
