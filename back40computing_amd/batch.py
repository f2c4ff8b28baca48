"""Python binding of batched BFS and the continuous-batching server
(C ABI in include/b40c_bfs_batch.h).  These have no counterpart in the
reference, which runs one search at a time."""
import ctypes

import numpy as np

from . import _native
from .bfs import BfsError, _check


def _destroy_quietly(obj):
    """destroy() from __del__: a partly built object, or an interpreter
    already tearing the module down, must not raise there."""
    try:
        if getattr(obj, "_h", None):
            obj.destroy()
    except Exception:
        pass


class GraphPool:
    """Device-resident graphs shared by batches and servers."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(_native.load().bfs_pool_create(device, ctypes.byref(self._h)), "bfs_pool_create")
        self.nodes = []
        self.edges = []

    def add(self, graph):
        gid = ctypes.c_int32()
        _check(_native.load().bfs_pool_add_graph(self._h, graph._s.row_offsets, graph._s.column_indices,
                                                  graph.nodes, graph.edges, ctypes.byref(gid)), "bfs_pool_add_graph")
        self.nodes.append(graph.nodes)
        self.edges.append(graph.edges)
        return gid.value

    def add_device(self, dgraph):
        """Adds a primitives.DeviceCsr (copied device to device)."""
        gid = ctypes.c_int32()
        _check(_native.load().bfs_pool_add_device_graph(self._h, ctypes.byref(dgraph._s), ctypes.byref(gid)),
               "bfs_pool_add_device_graph")
        self.nodes.append(dgraph.nodes)
        self.edges.append(dgraph.edges)
        return gid.value

    def device_bytes(self):
        b = ctypes.c_int64()
        _check(_native.load().bfs_pool_device_bytes(self._h, ctypes.byref(b)), "bfs_pool_device_bytes")
        return b.value

    def destroy(self):
        if self._h:
            _native.load().bfs_pool_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        _destroy_quietly(self)


def query_state_bytes(nodes):
    return int(_native.load().bfs_query_state_bytes(nodes))


class Batch:
    """Up to 256 queries advanced level by level together."""

    PERSISTENT, GRAPH = 0, 1

    def __init__(self, pool, max_queries=256):
        self.pool = pool
        self._h = ctypes.c_void_p()
        _check(_native.load().bfs_batch_create(pool._h, max_queries, ctypes.byref(self._h)), "bfs_batch_create")
        self.nq = 0
        self._graphs = None

    def set_config(self, **kw):
        cfg = _native.BfsConfig()
        _native.load().bfs_batch_config_default(ctypes.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, int(v))
        _check(_native.load().bfs_batch_set_config(self._h, ctypes.byref(cfg)), "bfs_batch_set_config")

    def enact(self, graph_ids, sources, direction_optimizing=True, mode=PERSISTENT):
        g = np.ascontiguousarray(graph_ids, dtype=np.int32)
        s = np.ascontiguousarray(sources, dtype=np.int32)
        st = _native.BatchStats()
        _check(_native.load().bfs_batch_enact(self._h, _native.i32ptr(g), _native.i32ptr(s), len(g),
                                               int(direction_optimizing), int(mode), ctypes.byref(st)),
               "bfs_batch_enact")
        self.nq = len(g)
        self._graphs = g
        return {f: getattr(st, f) for f, _ in st._fields_ if not f.startswith("reserved")}

    def latencies_ms(self):
        out = (ctypes.c_double * self.nq)()
        _check(_native.load().bfs_batch_query_latency(self._h, out), "bfs_batch_query_latency")
        return np.array(out[:], dtype=np.float64)

    def level_log(self):
        """[(top-down tiles, bottom-up tiles, start microseconds)] of the last enact."""
        td = (ctypes.c_uint32 * 64)()
        bu = (ctypes.c_uint32 * 64)()
        us = (ctypes.c_double * 64)()
        k = ctypes.c_int32()
        _check(_native.load().bfs_batch_level_log(self._h, 64, td, bu, us, ctypes.byref(k)), "bfs_batch_level_log")
        return [(int(td[i]), int(bu[i]), float(us[i])) for i in range(k.value)]

    def depths(self, q):
        out = np.empty(self.pool.nodes[int(self._graphs[q])], dtype=np.int32)
        _check(_native.load().bfs_batch_extract_depths(self._h, q, _native.i32ptr(out)), "bfs_batch_extract_depths")
        return out

    def destroy(self):
        if self._h:
            _native.load().bfs_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        _destroy_quietly(self)


class _PinnedBlock:
    """Page-locked host allocation, freed when the last array over it goes."""

    __slots__ = ("ptr",)

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        _check(_native.load().bfs_host_alloc_pinned(nbytes, ctypes.byref(self.ptr)), "bfs_host_alloc_pinned")

    def __del__(self):
        try:
            _native.load().bfs_host_free_pinned(self.ptr)
        except Exception:
            pass


def pinned_int32(n):
    """int32 numpy array of n elements in page-locked host memory."""
    blk = _PinnedBlock(max(n, 1) * 4)
    carr = (ctypes.c_int32 * max(n, 1)).from_address(blk.ptr.value)
    carr._owner = blk
    return np.ctypeslib.as_array(carr)[:n]


class Server:
    """Continuous-batching BFS server over a graph pool.

    Result buffers are page-locked host arrays from a per-graph free list:
    the server copies each finished query's labels into one on its own
    stream while later levels run.  wait() hands the array to the caller;
    recycle() returns it to the free list (optional)."""

    def __init__(self, pool, max_slots=64, levels_per_tick=2, hbm_budget_bytes=8 << 30, direction_optimizing=True):
        self.pool = pool
        self._free = {}              # graph id -> [array]
        cfg = _native.ServerConfig()
        cfg.max_slots = max_slots
        cfg.levels_per_tick = levels_per_tick
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.direction_optimizing = int(direction_optimizing)
        self._h = ctypes.c_void_p()
        _check(_native.load().bfs_server_create(pool._h, ctypes.byref(cfg), ctypes.byref(self._h)),
               "bfs_server_create")
        self._bufs = {}

    def _buffer(self, graph_id):
        free = self._free.setdefault(graph_id, [])
        try:
            return free.pop()
        except IndexError:
            return pinned_int32(self.pool.nodes[graph_id])

    def recycle(self, graph_id, buf):
        """Returns a result array from wait() to the free list."""
        self._free.setdefault(graph_id, []).append(buf)

    def submit(self, graph_id, src):
        """Returns a ticket; raises BfsError (status 8) when rejected."""
        buf = self._buffer(graph_id)
        t = ctypes.c_int64()
        rc = _native.load().bfs_server_submit(self._h, graph_id, int(src), _native.i32ptr(buf), ctypes.byref(t))
        if rc != 0:
            self.recycle(graph_id, buf)
            _check(rc, "bfs_server_submit")
        self._bufs[t.value] = buf
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """(depths, info dict) of a finished ticket."""
        info = _native.RequestInfo()
        rc = _native.load().bfs_server_wait(self._h, ticket, timeout_ms, ctypes.byref(info))
        if rc != 0:
            raise BfsError(rc, "bfs_server_wait")
        buf = self._bufs.pop(ticket)
        return buf, {f: getattr(info, f) for f, _ in info._fields_}

    def stats(self):
        st = _native.ServerStats()
        _check(_native.load().bfs_server_get_stats(self._h, ctypes.byref(st)), "bfs_server_get_stats")
        return {f: getattr(st, f) for f, _ in st._fields_}

    def destroy(self):
        """Finishes resident queries and joins the server threads; the
        result arrays of tickets not yet waited for stay valid after it."""
        if self._h:
            _native.load().bfs_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        # the native threads may still write into the pinned result arrays
        # of unfinished tickets: join them before those arrays can be freed
        _destroy_quietly(self)
