"""GPU parity of batched BFS and of the continuous-batching server.

Every query's depth labels must equal the serial oracle's, in both batch
modes (one persistent launch; hipGraph-captured per-level launches), with
and without direction optimization, and a query served alone must give the
same labels as the same query served in a mix.
"""
import threading
import time

import numpy as np
import pytest

import oracle
import back40computing_amd as b40c
from back40computing_amd import CsrGraph
from back40computing_amd.batch import Batch, GraphPool, Server, query_state_bytes
from back40computing_amd.workloads import batched_rmat, poisson_arrivals

pytestmark = pytest.mark.gpu


def mixed_graphs():
    return [CsrGraph.rmat(12, 8, seed=1), CsrGraph.rmat(14, 16, seed=2), CsrGraph.grid2d(60),
            CsrGraph.from_coo(6, [0, 0, 1, 2, 3, 4, 5], [1, 2, 3, 3, 4, 0, 0]),  # directed
            CsrGraph.erdos_renyi(3000, 9000, seed=3), CsrGraph.random(500, 900, seed=4),
            CsrGraph.from_arrays([0, 0], [])]  # single isolated vertex


@pytest.fixture(scope="module")
def mixed_pool(gpu):
    graphs = mixed_graphs()
    pool = GraphPool()
    for g in graphs:
        pool.add(g)
    yield graphs, pool
    pool.destroy()


def expected(g, src):
    return oracle.bfs(g.row_offsets, g.column_indices, int(src))[0]


@pytest.mark.parametrize("mode", [Batch.PERSISTENT, Batch.GRAPH], ids=["persistent", "graph"])
@pytest.mark.parametrize("direction", [True, False], ids=["do", "td"])
def test_batch_mixed_graphs(mixed_pool, mode, direction):
    graphs, pool = mixed_pool
    rng = np.random.default_rng(7 + mode)
    gids = rng.integers(0, len(graphs), 40).astype(np.int32)
    srcs = np.array([rng.integers(0, graphs[g].nodes) for g in gids], dtype=np.int32)
    b = Batch(pool, max_queries=64)
    st = b.enact(gids, srcs, direction_optimizing=direction, mode=mode)
    assert st["queries"] == 40
    for q in range(40):
        g = graphs[gids[q]]
        assert oracle.compare(b.depths(q), expected(g, srcs[q])) == -1, (q, gids[q], srcs[q])
    lat = b.latencies_ms()
    assert np.all(lat > 0) and np.all(lat <= st["elapsed_ms"] * 1.5 + 1.0)
    # the same batch again reuses the state arena and the captured graph
    st2 = b.enact(gids, srcs, direction_optimizing=direction, mode=mode)
    for q in range(0, 40, 7):
        assert oracle.compare(b.depths(q), expected(graphs[gids[q]], srcs[q])) == -1
    b.destroy()


def test_batch_switch_points(mixed_pool):
    graphs, pool = mixed_pool
    gids = np.array([0, 1, 2, 4, 1, 0], dtype=np.int32)
    srcs = np.array([graphs[g].sample_sources(1, seed=i)[0] for i, g in enumerate(gids)], dtype=np.int32)
    b = Batch(pool, max_queries=8)
    for alpha, beta in [(1000000, 1), (1, 24), (14, 1000000)]:
        b.set_config(do_alpha=alpha, do_beta=beta)
        for mode in (Batch.PERSISTENT, Batch.GRAPH):
            b.enact(gids, srcs, direction_optimizing=True, mode=mode)
            for q in range(len(gids)):
                assert oracle.compare(b.depths(q), expected(graphs[gids[q]], srcs[q])) == -1, (alpha, beta, q)
    b.destroy()


def test_batch_rejects_bad_arguments(mixed_pool):
    graphs, pool = mixed_pool
    b = Batch(pool, max_queries=4)
    with pytest.raises(b40c.BfsError):
        b.enact([0] * 5, [0] * 5)
    with pytest.raises(b40c.BfsError):
        b.enact([99], [0])
    with pytest.raises(b40c.BfsError):
        b.enact([0], [graphs[0].nodes])
    b.destroy()


def test_server_alone_and_in_mix(mixed_pool):
    graphs, pool = mixed_pool
    srv = Server(pool, max_slots=16, levels_per_tick=2, hbm_budget_bytes=256 << 20)
    probe = (1, int(graphs[1].sample_sources(1, seed=99)[0]))
    alone, info = srv.wait(srv.submit(*probe))
    assert info["status"] == 0 and info["finish_ms"] >= info["admit_ms"] >= info["submit_ms"]
    rng = np.random.default_rng(3)
    tickets = []
    for i in range(60):
        g = int(rng.integers(0, len(graphs)))
        s = int(rng.integers(0, graphs[g].nodes))
        tickets.append((g, s, srv.submit(g, s)))
        if i == 30:
            tickets.append((probe[0], probe[1], srv.submit(*probe)))
    for g, s, t in tickets:
        d, info = srv.wait(t)
        assert oracle.compare(d, expected(graphs[g], s)) == -1, (g, s)
        if (g, s) == probe:
            assert np.array_equal(d, alone)
    st = srv.stats()
    assert st["completed"] == 62 and st["max_resident"] > 1
    srv.destroy()


def test_server_admission_budget(mixed_pool):
    graphs, pool = mixed_pool
    # budget for about two queries on the largest graph: the rest must queue
    big = max(range(len(graphs)), key=lambda i: graphs[i].nodes)
    budget = 2 * query_state_bytes(graphs[big].nodes) + 4096
    srv = Server(pool, max_slots=32, hbm_budget_bytes=budget)
    tickets = [(s, srv.submit(big, s)) for s in graphs[big].sample_sources(6, seed=5)]
    for s, t in tickets:
        d, _ = srv.wait(t)
        assert oracle.compare(d, expected(graphs[big], s)) == -1
    assert srv.stats()["max_resident"] <= 2
    srv.destroy()
    tiny = Server(pool, max_slots=4, hbm_budget_bytes=1024)
    with pytest.raises(b40c.BfsError) as ei:
        tiny.submit(big, 0)
    assert ei.value.status == 8  # BFS_ERR_REJECTED
    assert tiny.stats()["rejected"] == 1
    tiny.destroy()


def test_server_concurrent_submitters(mixed_pool):
    graphs, pool = mixed_pool
    srv = Server(pool, max_slots=8, levels_per_tick=1, hbm_budget_bytes=128 << 20)
    errors = []
    arrivals = poisson_arrivals(2000.0, 40, seed=11)

    def client(k):
        rng = np.random.default_rng(100 + k)
        t0 = time.perf_counter()
        for i in range(k, 40, 4):
            while time.perf_counter() - t0 < arrivals[i]:
                time.sleep(0.0001)
            g = int(rng.integers(0, len(graphs)))
            s = int(rng.integers(0, graphs[g].nodes))
            d, _ = srv.wait(srv.submit(g, s))
            if oracle.compare(d, expected(graphs[g], s)) != -1:
                errors.append((g, s))

    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors
    assert srv.stats()["completed"] == 40
    srv.destroy()


@pytest.mark.slow
def test_config4_batched_rmat_pool(gpu):
    specs, graphs, gids, srcs = batched_rmat()
    pool = GraphPool()
    for g in graphs:
        pool.add(g)
    b = Batch(pool, max_queries=256)
    for mode in (Batch.PERSISTENT, Batch.GRAPH):
        b.enact(gids, srcs, direction_optimizing=True, mode=mode)
        for q in range(0, 256, 1 if mode == Batch.PERSISTENT else 5):
            assert oracle.compare(b.depths(q), expected(graphs[gids[q]], srcs[q])) == -1, (mode, q)
    b.destroy()
    pool.destroy()
