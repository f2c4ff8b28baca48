"""CSR construction on the device (b40c_device_csr_*) against the host
builders of graph_builder.cpp: the same input gives the same row offsets and
column indices, bit for bit (rows sorted by column on both sides).  The
degree sum (row[V] == E) and the transpose's own transpose close the loop."""
import numpy as np
import pytest

import oracle
import back40computing_amd as b40c
from back40computing_amd import CsrGraph
from back40computing_amd.batch import Batch, GraphPool
from back40computing_amd.primitives import DeviceCsr

pytestmark = pytest.mark.gpu


def same_csr(host, dev):
    h = dev.to_host()
    assert h.nodes == host.nodes and h.edges == host.edges
    assert np.array_equal(h.row_offsets, host.row_offsets)
    assert np.array_equal(h.column_indices, host.column_indices)
    return h


@pytest.mark.parametrize("scale,ef,scramble", [(10, 16, True), (14, 8, False), (16, 16, True), (18, 32, True)])
def test_rmat_matches_host_builder(gpu, scale, ef, scramble):
    host = CsrGraph.rmat(scale, ef, seed=0x77 + scale, scramble=scramble)
    dev = DeviceCsr.rmat(scale, ef, seed=0x77 + scale, scramble=scramble)
    assert dev.symmetric
    same_csr(host, dev)
    dev.free()


@pytest.mark.parametrize("flags", range(8))
def test_from_coo_matches_host_builder(gpu, flags):
    sym, dedup, loops = bool(flags & 1), bool(flags & 2), bool(flags & 4)
    rng = np.random.default_rng(100 + flags)
    for nodes, m in [(1, 0), (1, 5), (7, 40), (1000, 0), (1000, 5000), (40000, 300000), (70001, 2_000_000)]:
        src = rng.integers(0, nodes, m).astype(np.int32)
        dst = rng.integers(0, nodes, m).astype(np.int32)
        if m:
            dst[: m // 10] = src[: m // 10]                  # some self-loops
            src[m // 2: m // 2 + m // 10] = src[: m // 10]   # and repeats
            dst[m // 2: m // 2 + m // 10] = dst[: m // 10]
        host = CsrGraph.from_coo(nodes, src, dst, symmetrize=sym, dedup=dedup, drop_self_loops=loops)
        dev = DeviceCsr.from_coo(nodes, src, dst, symmetrize=sym, dedup=dedup, drop_self_loops=loops)
        same_csr(host, dev)
        assert dev.symmetric == sym
        dev.free()


def test_transpose_matches_host(gpu):
    rng = np.random.default_rng(7)
    nodes, m = 5000, 40000
    src = rng.integers(0, nodes, m).astype(np.int32)
    dst = rng.integers(0, nodes, m).astype(np.int32)
    dev = DeviceCsr.from_coo(nodes, src, dst)
    t = dev.transpose()
    host_t = CsrGraph.from_coo(nodes, dst, src)
    same_csr(host_t, t)
    tt = t.transpose()
    same_csr(CsrGraph.from_coo(nodes, src, dst), tt)


def test_pool_search_on_device_built_graphs(gpu):
    # one symmetric R-MAT and one directed random graph (in-neighbour CSR
    # built on the device): batched depths against the oracle
    g1 = DeviceCsr.rmat(16, 16, seed=5)
    rng = np.random.default_rng(9)
    n2, m2 = 30000, 120000
    s2, d2 = rng.integers(0, n2, m2).astype(np.int32), rng.integers(0, n2, m2).astype(np.int32)
    g2 = DeviceCsr.from_coo(n2, s2, d2)
    hosts = [g1.to_host(), g2.to_host()]
    pool = GraphPool()
    ids = [pool.add_device(g1), pool.add_device(g2)]
    qs = [(ids[0], int(hosts[0].sample_sources(1, seed=1)[0])), (ids[1], int(hosts[1].sample_sources(1, seed=2)[0]))]
    b = Batch(pool, max_queries=4)
    for do in (True, False):
        b.enact([q[0] for q in qs], [q[1] for q in qs], direction_optimizing=do)
        for k, (gid, src) in enumerate(qs):
            expect, _, _ = oracle.bfs(hosts[gid].row_offsets, hosts[gid].column_indices, src)
            assert oracle.compare(b.depths(k), expect) == -1, (gid, do)
    b.destroy()
    pool.destroy()


@pytest.mark.slow
def test_rmat24_device_build(gpu):
    # the serving pool's largest graph, built on the device; spot checks
    # against the host builder's degree sequence and first rows
    dev = DeviceCsr.rmat(24, 16, seed=0x5E4E + 1000)
    print("R-MAT scale 24 built on the device in %.1f ms: V=%d E=%d" % (dev.build_ms, dev.nodes, dev.edges))
    host = CsrGraph.rmat(24, 16, seed=0x5E4E + 1000)
    same_csr(host, dev)
    dev.free()
