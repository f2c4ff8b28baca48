"""Graph files through the whole path: a DIMACS, METIS or Matrix Market file
read by the library's readers, its CSR built on the device (sorted,
deduplicated, loops dropped) or on the host, then searched -- single-source
strategies on the host-built CSR and a batch on the device-built one --
against the oracle run on the host-built CSR."""
import numpy as np
import pytest

import oracle
import back40computing_amd as b40c
from back40computing_amd import CsrGraph, CsrProblem, read_coo
from back40computing_amd.batch import Batch, GraphPool
from back40computing_amd.primitives import DeviceCsr

pytestmark = pytest.mark.gpu


def write_files(tmp_path, n, src, dst):
    """The directed pairs (src[i], dst[i]) in each format (1-based ids)."""
    d = tmp_path / "g.gr"
    d.write_text("p sp %d %d\n" % (n, len(src)) + "".join("a %d %d 1\n" % (u + 1, v + 1) for u, v in zip(src, dst)))
    # METIS lists each vertex's neighbours on its own line
    order = np.argsort(src, kind="stable")
    rows = [[] for _ in range(n)]
    for i in order:
        rows[src[i]].append(dst[i] + 1)
    mt = tmp_path / "g.graph"
    mt.write_text("%d %d\n" % (n, len(src)) + "".join(" ".join(map(str, r)) + "\n" for r in rows))
    # Matrix Market: the entry "i j" is the arc j -> i
    mm = tmp_path / "g.mtx"
    mm.write_text("%%MatrixMarket matrix coordinate pattern general\n%d %d %d\n" % (n, n, len(src)) +
                  "".join("%d %d\n" % (v + 1, u + 1) for u, v in zip(src, dst)))
    return {"dimacs": str(d), "metis": str(mt), "market": str(mm)}


@pytest.mark.parametrize("kind", ["dimacs", "metis", "market"])
def test_file_to_device_csr_to_search(gpu, tmp_path, kind):
    rng = np.random.default_rng(0xF11E)
    n, m = 3000, 15000
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    path = write_files(tmp_path, n, src, dst)[kind]
    nodes, s, d = read_coo(kind, path)
    assert nodes == n and len(s) == m
    host = CsrGraph.from_coo(nodes, s, d, dedup=True, drop_self_loops=True)
    dev = DeviceCsr.from_coo(nodes, s, d, dedup=True, drop_self_loops=True)
    back = dev.to_host()
    assert np.array_equal(back.row_offsets, host.row_offsets)
    assert np.array_equal(back.column_indices, host.column_indices)
    row, col = host.row_offsets, host.column_indices
    sources = host.sample_sources(3, seed=9)
    expects = [oracle.bfs(row, col, int(x))[0] for x in sources]
    # single-source strategies on the host-built CSR (directed: bottom-up
    # levels use the in-neighbour CSR)
    p = CsrProblem(mark_predecessors=True).from_host_problem(host)
    for strategy, fused in [(b40c.EXPAND_CONTRACT, True), (b40c.CONTRACT_EXPAND, True), (b40c.TWO_PHASE, False),
                            (b40c.HYBRID, True), (b40c.DIRECTION_OPTIMIZING, True)]:
        for x, expect in zip(sources, expects):
            ft = b40c.VERTEX_FRONTIERS if strategy in (b40c.EXPAND_CONTRACT, b40c.DIRECTION_OPTIMIZING) \
                else b40c.EDGE_FRONTIERS
            p.reset(ft, 1.3)
            p.enact(strategy, int(x), fused=fused)
            assert oracle.compare(p.extract_depths(), expect) == -1, (kind, strategy, int(x))
            assert oracle.check_preds(row, col, expect, p.extract_results()) == -1, (kind, strategy, int(x))
    p.destroy()
    # a batch on the device-built CSR (added to the pool without a host copy)
    pool = GraphPool()
    gid = pool.add_device(dev)
    b = Batch(pool, max_queries=4)
    for do in (True, False):
        b.enact([gid] * len(sources), [int(x) for x in sources], direction_optimizing=do)
        for q, expect in enumerate(expects):
            assert oracle.compare(b.depths(q), expect) == -1, (kind, do, q)
    b.destroy()
    pool.destroy()
    dev.free()
