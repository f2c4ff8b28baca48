"""Search statistics (Enactor*::GetStatistics, e.g.
b40c/graph/bfs/enactor_contract_expand.cuh:236-250) against invariants of the
traversal: total_queued sums the input length of every step, so with exact
claims each strategy's count follows from the visited vertices V_vis and
their edges E_vis (counted as test/bfs/test_bfs.cu:360-367); and the
instrumented duty (b40c/util/kernel_runtime_stats.cuh:242-295) is a fraction
of each workgroup's lifetime."""
import numpy as np
import pytest

import oracle
import back40computing_amd as b40c
from back40computing_amd import CsrGraph, CsrProblem

pytestmark = pytest.mark.gpu


def graphs():
    return [CsrGraph.rmat(15, 16, seed=0x51), CsrGraph.grid2d(60), CsrGraph.erdos_renyi(20000, 80000, seed=3)]


def visited(g, src):
    expect, levels, _ = oracle.bfs(g.row_offsets, g.column_indices, int(src))
    reached = expect >= 0
    return int(reached.sum()), int(g.degrees()[reached].sum()), levels


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "iter"])
def test_total_queued_with_exact_claims(gpu, fused):
    for g in graphs():
        p = CsrProblem().from_host_problem(g)
        for src in g.sample_sources(2, seed=5):
            nv, ev, levels = visited(g, src)
            for strategy in (b40c.EXPAND_CONTRACT, b40c.DIRECTION_OPTIMIZING):
                p.reset(b40c.VERTEX_FRONTIERS, 1.0)
                st = p.enact(strategy, int(src), fused=fused)
                # every visited vertex is in exactly one vertex frontier
                assert st.total_queued == nv, (strategy, st.total_queued, nv)
                assert st.search_depth == levels
            p.reset(b40c.EDGE_FRONTIERS, 1.3)
            st = p.enact(b40c.CONTRACT_EXPAND, int(src), fused=fused)
            # the source, then every neighbour of every claimed vertex
            assert st.total_queued == 1 + ev, (st.total_queued, ev)
            st = p.enact(b40c.TWO_PHASE, int(src), fused=fused)
            # contract + expand + filter inputs: each level's claimed
            # vertices, all their edges, and the filtered edges (at least one
            # per vertex discovered after the source, at most all edges)
            assert 2 * nv + ev <= st.total_queued <= 1 + nv + 2 * ev, (st.total_queued, nv, ev)
            st = p.enact(b40c.HYBRID, int(src))
            assert st.total_queued >= nv
        p.destroy()


def test_total_queued_best_effort_is_at_least_exact(gpu):
    g = CsrGraph.rmat(15, 16, seed=0x52)
    p = CsrProblem().from_host_problem(g)
    p.set_config(cull_mode=b40c.CULL_BEST_EFFORT)
    for src in g.sample_sources(2, seed=6):
        nv, ev, _ = visited(g, src)
        p.reset(b40c.VERTEX_FRONTIERS, 4.0)
        assert p.enact(b40c.EXPAND_CONTRACT, int(src)).total_queued >= nv
        p.reset(b40c.EDGE_FRONTIERS, 64.0)
        assert p.enact(b40c.CONTRACT_EXPAND, int(src)).total_queued >= 1 + ev
    p.destroy()


def test_instrumented_duty_is_a_fraction(gpu):
    g = CsrGraph.rmat(16, 16, seed=0x53)
    p = CsrProblem().from_host_problem(g)
    src = int(g.sample_sources(1, seed=7)[0])
    for strategy in (b40c.EXPAND_CONTRACT, b40c.CONTRACT_EXPAND, b40c.TWO_PHASE, b40c.HYBRID,
                     b40c.DIRECTION_OPTIMIZING):
        p.set_config(instrument=1)
        ft = b40c.VERTEX_FRONTIERS if strategy in (b40c.EXPAND_CONTRACT, b40c.DIRECTION_OPTIMIZING) \
            else b40c.EDGE_FRONTIERS
        p.reset(ft, 1.3)
        st = p.enact(strategy, src)
        assert 0.0 < st.avg_duty <= 1.0, (strategy, st.avg_duty)
        p.set_config(instrument=0)
        p.reset(ft, 1.3)
        assert p.enact(strategy, src).avg_duty == 0.0
    p.destroy()
