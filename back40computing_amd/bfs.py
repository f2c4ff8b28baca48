"""Python mirror of the reference BFS interface over the C ABI.

Names follow the reference's classes so that tests read like its driver
(test/bfs/test_bfs.cu:584-783):

    CsrGraph                 <- b40c/graph/csr_graph.cuh:50 (+ builder/*.cuh generators)
    CsrProblem               <- b40c/graph/bfs/csr_problem.cuh:68
        .from_host_problem   <- FromHostProblem  (csr_problem.cuh:328)
        .reset               <- Reset            (csr_problem.cuh:547)
        .extract_results     <- ExtractResults   (csr_problem.cuh:255)
    EnactorExpandContract    <- b40c/graph/bfs/enactor_expand_contract.cuh:58
    EnactorContractExpand    <- b40c/graph/bfs/enactor_contract_expand.cuh:60
    EnactorTwoPhase          <- b40c/graph/bfs/enactor_two_phase.cuh:67
    EnactorHybrid            <- b40c/graph/bfs/enactor_hybrid.cuh:63

Every call goes through libb40c_hip.so; errors raise BfsError carrying the
library's status code (the reference returned cudaError_t).
"""
import ctypes

import numpy as np

from . import _native

EXPAND_CONTRACT, CONTRACT_EXPAND, TWO_PHASE, HYBRID, DIRECTION_OPTIMIZING = 0, 1, 2, 3, 4
VERTEX_FRONTIERS, EDGE_FRONTIERS, MIXED_FRONTIERS = 0, 1, 2
CULL_EXACT, CULL_BEST_EFFORT = 0, 1

STATUS = {
    0: "BFS_OK", 1: "BFS_ERR_INVALID_ARG", 2: "BFS_ERR_OOM", 3: "BFS_ERR_OVERFLOW",
    4: "BFS_ERR_UNSUPPORTED", 5: "BFS_ERR_LAUNCH", 6: "BFS_ERR_HIP", 7: "BFS_ERR_NO_DEVICE",
    8: "BFS_ERR_REJECTED", 9: "BFS_ERR_TIMEOUT", 10: "BFS_ERR_NOT_READY",
}


class BfsError(RuntimeError):
    def __init__(self, status, what):
        lib = _native.load()
        msg = lib.bfs_last_error_message().decode() or lib.bfs_status_string(status).decode()
        super().__init__("%s: %s (%s)" % (what, msg, STATUS.get(status, status)))
        self.status = status


def _check(status, what):
    if status != 0:
        raise BfsError(status, what)


class CsrGraph:
    """Host CSR graph (int32 row offsets and column indices).

    Built by the library's generators (memory owned by the library) or
    wrapped from numpy arrays (memory owned by the arrays).
    """

    def __init__(self):
        self._s = _native.CsrStruct()
        self._owned = False
        self._keep = None

    # -- construction -----------------------------------------------------
    @classmethod
    def _build(cls, fn, *args):
        g = cls()
        rc = fn(*args, ctypes.byref(g._s))
        if rc != 0:
            raise BfsError(1, "graph builder failed with code %d" % rc)
        g._owned = True
        return g

    @classmethod
    def grid2d(cls, width):
        return cls._build(_native.load().b40c_graph_grid2d, width)

    @classmethod
    def grid3d(cls, width):
        return cls._build(_native.load().b40c_graph_grid3d, width)

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=0x5EED0022, scramble=True):
        return cls._build(_native.load().b40c_graph_rmat, scale, edge_factor, a, b, c, seed, int(scramble))

    @classmethod
    def erdos_renyi(cls, n, m, seed=0xE5D05):
        return cls._build(_native.load().b40c_graph_erdos_renyi, n, m, seed)

    @classmethod
    def random(cls, n, m, seed=0, undirected=False):
        return cls._build(_native.load().b40c_graph_random, n, m, seed, int(undirected))

    @classmethod
    def from_coo(cls, nodes, src, dst, symmetrize=False, dedup=False, drop_self_loops=False):
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        assert src.shape == dst.shape
        return cls._build(_native.load().b40c_graph_from_coo, nodes, len(src), _native.i32ptr(src),
                          _native.i32ptr(dst), int(symmetrize), int(dedup), int(drop_self_loops))

    @classmethod
    def from_arrays(cls, row_offsets, column_indices):
        row = np.ascontiguousarray(row_offsets, dtype=np.int32)
        col = np.ascontiguousarray(column_indices, dtype=np.int32)
        g = cls()
        g._keep = (row, col)
        g._s.nodes = len(row) - 1
        g._s.edges = len(col)
        g._s.row_offsets = _native.i32ptr(row)
        g._s.column_indices = _native.i32ptr(col) if len(col) else _native.i32ptr(np.zeros(1, np.int32))
        return g

    @classmethod
    def from_file(cls, kind, path, undirected=False):
        """CSR of a DIMACS / METIS / Matrix Market file (read_coo), with
        every pair kept (no dedup, no self-loop removal)."""
        nodes, src, dst = read_coo(kind, path, undirected)
        return cls.from_coo(nodes, src, dst)

    # -- access -----------------------------------------------------------
    @property
    def nodes(self):
        return int(self._s.nodes)

    @property
    def edges(self):
        return int(self._s.edges)

    @property
    def row_offsets(self):
        return np.ctypeslib.as_array(self._s.row_offsets, shape=(self.nodes + 1,))

    @property
    def column_indices(self):
        if self.edges == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(self._s.column_indices, shape=(self.edges,))

    def degrees(self):
        r = self.row_offsets
        return (r[1:] - r[:-1]).astype(np.int64)

    def sample_sources(self, count, seed=0x50C4CE5):
        out = np.zeros(count, dtype=np.int32)
        rc = _native.load().b40c_graph_sample_sources(ctypes.byref(self._s), count, seed, _native.i32ptr(out))
        if rc != 0:
            raise BfsError(1, "sample_sources failed")
        return out

    def free(self):
        if self._owned:
            _native.load().b40c_graph_free(ctypes.byref(self._s))
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def read_coo(kind, path, undirected=False):
    """(nodes, src, dst) of a graph file: kind is "dimacs", "metis" or
    "market" (b40c/graph/builder/{dimacs,metis,market}.cuh grammars)."""
    lib = _native.load()
    c = _native.CooStruct()
    p = path.encode() if isinstance(path, str) else path
    if kind == "dimacs":
        rc = lib.b40c_read_dimacs(p, int(undirected), ctypes.byref(c))
    elif kind == "metis":
        rc = lib.b40c_read_metis(p, ctypes.byref(c))
    elif kind == "market":
        rc = lib.b40c_read_market(p, int(undirected), ctypes.byref(c))
    else:
        raise ValueError("unknown graph file kind %r" % kind)
    if rc != 0:
        raise BfsError(1, "reading %s file %s failed with code %d" % (kind, path, rc))
    try:
        n = int(c.pairs)
        src = np.ctypeslib.as_array(c.src, shape=(max(n, 1),))[:n].copy()
        dst = np.ctypeslib.as_array(c.dst, shape=(max(n, 1),))[:n].copy()
        return int(c.nodes), src, dst
    finally:
        lib.b40c_coo_free(ctypes.byref(c))


class Stats:
    def __init__(self, s):
        self.total_queued = int(s.total_queued)
        self.search_depth = int(s.search_depth)
        self.avg_duty = float(s.avg_duty)
        self.elapsed_ms = float(s.elapsed_ms)
        self.kernel_launches = int(s.kernel_launches)

    def __repr__(self):
        return ("Stats(elapsed_ms=%.4f, search_depth=%d, total_queued=%d, launches=%d)"
                % (self.elapsed_ms, self.search_depth, self.total_queued, self.kernel_launches))


class CsrProblem:
    """Device-resident BFS problem (CSR, labels, visited mask, frontier queues)."""

    def __init__(self, mark_predecessors=False):
        self.mark_predecessors = bool(mark_predecessors)
        self._h = ctypes.c_void_p()
        self.nodes = 0
        self.edges = 0

    def from_host_problem(self, graph, device=0):
        lib = _native.load()
        _check(lib.bfs_problem_create(graph._s.row_offsets, graph._s.column_indices, graph.nodes,
                                      graph.edges, int(self.mark_predecessors), device,
                                      ctypes.byref(self._h)), "bfs_problem_create")
        self.nodes = graph.nodes
        self.edges = graph.edges
        return self

    def reset(self, frontier_type, queue_sizing=1.3):
        _check(_native.load().bfs_problem_reset(self._h, frontier_type, queue_sizing), "bfs_problem_reset")

    def set_config(self, **kw):
        lib = _native.load()
        cfg = _native.BfsConfig()
        _check(lib.bfs_problem_get_config(self._h, ctypes.byref(cfg)), "get_config")
        for k, v in kw.items():
            setattr(cfg, k, int(v))
        _check(lib.bfs_problem_set_config(self._h, ctypes.byref(cfg)), "set_config")

    def get_config(self):
        cfg = _native.BfsConfig()
        _check(_native.load().bfs_problem_get_config(self._h, ctypes.byref(cfg)), "get_config")
        return {f: getattr(cfg, f) for f, _ in _native.BfsConfig._fields_}

    def enact(self, strategy, src, fused=True, max_grid_size=0):
        st = _native.BfsStats()
        _check(_native.load().bfs_enact(self._h, strategy, int(fused), int(src), int(max_grid_size),
                                         ctypes.byref(st)), "bfs_enact")
        return Stats(st)

    def extract_results(self):
        out = np.empty(self.nodes, dtype=np.int32)
        _check(_native.load().bfs_extract_labels(self._h, _native.i32ptr(out)), "bfs_extract_labels")
        return out

    def extract_depths(self):
        out = np.empty(self.nodes, dtype=np.int32)
        _check(_native.load().bfs_extract_depths(self._h, _native.i32ptr(out)), "bfs_extract_depths")
        return out

    def level_log(self):
        """[(input length, start microseconds)] of each step of the last enact."""
        lens = (ctypes.c_uint32 * 64)()
        us = (ctypes.c_double * 64)()
        k = ctypes.c_int32()
        _check(_native.load().bfs_get_level_log(self._h, 64, lens, us, ctypes.byref(k)), "level_log")
        return [(int(lens[i]), float(us[i])) for i in range(k.value)]

    def device_info(self):
        cus, grid, nbytes = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _check(_native.load().bfs_problem_device_info(self._h, ctypes.byref(cus), ctypes.byref(grid),
                                                       ctypes.byref(nbytes)), "device_info")
        return {"num_cus": cus.value, "grid_size": grid.value, "device_bytes": nbytes.value}

    def destroy(self):
        if self._h:
            _native.load().bfs_problem_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class _Enactor:
    STRATEGY = None
    FUSED = True

    def __init__(self, verbose=False):
        self.verbose = verbose
        self.last = None

    def get_frontier_type(self):
        return _native.load().bfs_strategy_frontier_type(self.STRATEGY)

    def _enact(self, problem, src, fused, max_grid_size):
        self.last = problem.enact(self.STRATEGY, src, fused=fused, max_grid_size=max_grid_size)
        return self.last

    def get_statistics(self):
        """(total_queued, search_depth, avg_duty), as Enactor*::GetStatistics."""
        s = self.last
        return (s.total_queued, s.search_depth, s.avg_duty) if s else (0, 0, 0.0)


class EnactorExpandContract(_Enactor):
    STRATEGY = EXPAND_CONTRACT

    def enact_fused_search(self, problem, src, max_grid_size=0):
        return self._enact(problem, src, True, max_grid_size)

    def enact_iterative_search(self, problem, src, max_grid_size=0):
        return self._enact(problem, src, False, max_grid_size)


class EnactorContractExpand(EnactorExpandContract):
    STRATEGY = CONTRACT_EXPAND


class EnactorTwoPhase(EnactorExpandContract):
    STRATEGY = TWO_PHASE


class EnactorDirectionOptimizing(EnactorExpandContract):
    """Addition: top-down expand-contract levels mixed with bottom-up levels."""
    STRATEGY = DIRECTION_OPTIMIZING


class EnactorHybrid(_Enactor):
    STRATEGY = HYBRID

    def enact_search(self, problem, src, max_grid_size=0):
        return self._enact(problem, src, True, max_grid_size)


def device_count():
    n = ctypes.c_int32()
    _check(_native.load().bfs_device_count(ctypes.byref(n)), "bfs_device_count")
    return n.value


def barrier_benchmark(blocks_per_cu=1, iterations=10000, device=0):
    us, grid = ctypes.c_double(), ctypes.c_int32()
    _check(_native.load().bfs_barrier_benchmark(device, blocks_per_cu, iterations, ctypes.byref(us),
                                                 ctypes.byref(grid)), "bfs_barrier_benchmark")
    return us.value, grid.value


def selftest_wave_primitives(device=0):
    bad = ctypes.c_int32()
    rc = _native.load().bfs_selftest_wave_primitives(device, ctypes.byref(bad))
    return rc, bad.value
