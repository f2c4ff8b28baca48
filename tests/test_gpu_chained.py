"""Chained (single-pass, decoupled look-back) compaction of the contract and
filter steps (config compaction = 2) against the oracle: two-phase fused and
per-level, and the hybrid forced through its two-phase levels, in both cull
modes, depths and predecessors; on graphs whose steps span one tile, a few
tiles and thousands of tiles, and grids of one and three workgroups (a
workgroup then owns many consecutive tiles and continues its own prefix)."""
import numpy as np
import pytest

import oracle
import back40computing_amd as b40c
from back40computing_amd import CsrGraph, CsrProblem

pytestmark = pytest.mark.gpu

CHAINED = [(b40c.TWO_PHASE, True, {}), (b40c.TWO_PHASE, False, {}),
           (b40c.HYBRID, True, {"saturation_threshold": 1}), (b40c.HYBRID, True, {"saturation_threshold": 300})]
CHAINED_IDS = ["tp_fused", "tp_iter", "hybrid_sat1", "hybrid_sat300"]


def check(g, p, strategy, fused, src, max_grid_size=0):
    row, col = g.row_offsets, g.column_indices
    expect, levels, _ = oracle.bfs(row, col, int(src))
    p.reset(b40c.EDGE_FRONTIERS, 1.3 if p.get_config()["cull_mode"] == b40c.CULL_EXACT else 64.0)
    st = p.enact(strategy, int(src), fused=fused, max_grid_size=max_grid_size)
    got = p.extract_depths()
    bad = oracle.compare(got, expect)
    assert bad == -1, (strategy, fused, int(src), bad)
    assert st.search_depth == levels
    if p.mark_predecessors:
        assert oracle.check_preds(row, col, expect, p.extract_results()) == -1
    return st


@pytest.mark.parametrize("variant", range(len(CHAINED)), ids=CHAINED_IDS)
@pytest.mark.parametrize("cull", [b40c.CULL_EXACT, b40c.CULL_BEST_EFFORT], ids=["exact", "best_effort"])
@pytest.mark.parametrize("mark_pred", [False, True], ids=["depth", "pred"])
def test_chained_compaction_parity(gpu, variant, cull, mark_pred):
    strategy, fused, extra = CHAINED[variant]
    graphs = [CsrGraph.rmat(14, 16, seed=5), CsrGraph.grid2d(40), CsrGraph.random(3000, 20000, seed=8)]
    for g in graphs:
        p = CsrProblem(mark_predecessors=mark_pred).from_host_problem(g)
        p.set_config(compaction=2, cull_mode=cull, **extra)
        for src in g.sample_sources(2, seed=3):
            for grid in (0, 1, 3):
                check(g, p, strategy, fused, src, max_grid_size=grid)
        p.destroy()


def test_chained_matches_atomic_statistics(gpu):
    """With exact claims the queued totals of both compaction modes are the
    same numbers (each step's output is a set fixed by the level); only the
    order within a queue differs."""
    g = CsrGraph.rmat(16, 16, seed=9)
    p = CsrProblem().from_host_problem(g)
    for src in g.sample_sources(3, seed=2):
        tq = []
        for mode in (1, 2):
            p.set_config(compaction=mode)
            for fused in (True, False):
                st = check(g, p, b40c.TWO_PHASE, fused, src)
                tq.append(st.total_queued)
        assert len(set(tq)) == 1, tq
    p.destroy()


def test_chained_overflow_is_reported(gpu):
    g = CsrGraph.rmat(12, 16, seed=4)
    p = CsrProblem().from_host_problem(g)
    p.set_config(compaction=2)
    src = int(g.sample_sources(1, seed=1)[0])
    p.reset(b40c.EDGE_FRONTIERS, 0.01)
    with pytest.raises(b40c.BfsError) as e:
        p.enact(b40c.TWO_PHASE, src)
    assert e.value.status == 3  # BFS_ERR_OVERFLOW
    check(g, p, b40c.TWO_PHASE, True, src)  # recovers after a reset with room
    p.destroy()


def test_chained_many_searches_wrap_epochs(gpu):
    """Each search advances the status-word epoch by 3 (V + 2); with 2^24
    vertices (a scale-12 R-MAT component plus isolated vertices) the 30-bit
    epoch space is used up after about 19 searches and the words are cleared:
    searches on both sides of the clear stay exact."""
    small = CsrGraph.rmat(12, 16, seed=12)
    row, col = small.row_offsets, small.column_indices
    src = np.repeat(np.arange(small.nodes, dtype=np.int32), np.diff(row))
    g = CsrGraph.from_coo(1 << 24, src, col)
    p = CsrProblem().from_host_problem(g)
    p.set_config(compaction=2)
    p.reset(b40c.EDGE_FRONTIERS, 1.3)
    sources = small.sample_sources(4, seed=6)
    expects = [oracle.bfs(g.row_offsets, g.column_indices, int(s))[0] for s in sources]
    for it in range(26):
        k = it % len(sources)
        p.enact(b40c.TWO_PHASE, int(sources[k]), fused=it % 3 != 0)
        assert oracle.compare(p.extract_depths(), expects[k]) == -1, it
    p.destroy()
