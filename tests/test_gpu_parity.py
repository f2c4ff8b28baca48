"""GPU parity: every strategy's depth labels are bit-exact against the CPU oracle.

All calls go through the C ABI (libb40c_hip.so).  Predecessors are checked by
the invariant of test/bfs/test_bfs.cu:387-433 (tightened), since racing
discoverers may legitimately pick different parents.
"""
import json
import os

import numpy as np
import pytest

import oracle
import back40computing_amd as b40c
from back40computing_amd import CsrGraph, CsrProblem

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tiny_graphs.json")

# (strategy, fused) pairs: the reference's four enactors, persistent and per-level.
VARIANTS = [
    (b40c.EXPAND_CONTRACT, True), (b40c.EXPAND_CONTRACT, False),
    (b40c.CONTRACT_EXPAND, True), (b40c.CONTRACT_EXPAND, False),
    (b40c.TWO_PHASE, True), (b40c.TWO_PHASE, False),
    (b40c.HYBRID, True),
    (b40c.DIRECTION_OPTIMIZING, True), (b40c.DIRECTION_OPTIMIZING, False),
]
VARIANT_IDS = ["ec_fused", "ec_iter", "ce_fused", "ce_iter", "tp_fused", "tp_iter", "hybrid", "do_fused", "do_iter"]


def frontier_type(strategy):
    if strategy in (b40c.EXPAND_CONTRACT, b40c.DIRECTION_OPTIMIZING):
        return b40c.VERTEX_FRONTIERS
    return b40c.EDGE_FRONTIERS


def run(problem, strategy, fused, src, queue_sizing=1.3, max_grid_size=0):
    problem.reset(frontier_type(strategy), queue_sizing)
    stats = problem.enact(strategy, int(src), fused=fused, max_grid_size=max_grid_size)
    return problem.extract_depths(), stats


def assert_parity(g, problem, strategy, fused, src, **kw):
    row, col = g.row_offsets, g.column_indices
    expect, levels, _ = oracle.bfs(row, col, int(src))
    got, stats = run(problem, strategy, fused, src, **kw)
    bad = oracle.compare(got, expect)
    assert bad == -1, "vertex %d: got %d expected %d (strategy %d fused %d src %d)" % (
        bad, got[bad], expect[bad], strategy, fused, src)
    assert stats.search_depth == levels
    if problem.mark_predecessors:
        preds = problem.extract_results()
        assert oracle.check_preds(row, col, expect, preds) == -1
    return stats


def test_wave_primitives_selftest(gpu):
    rc, bad = b40c.selftest_wave_primitives()
    assert rc == 0 and bad == 0, bad


def test_grid_barrier_runs(gpu):
    us, grid = b40c.barrier_benchmark(blocks_per_cu=2, iterations=2000)
    assert grid > 0 and 0 < us < 1000


@pytest.mark.parametrize("variant", range(len(VARIANTS)), ids=VARIANT_IDS)
@pytest.mark.parametrize("mark_pred", [False, True], ids=["depth", "pred"])
def test_golden_graphs(gpu, variant, mark_pred):
    strategy, fused = VARIANTS[variant]
    with open(GOLDEN) as f:
        graphs = json.load(f)["graphs"]
    for gd in graphs:
        g = CsrGraph.from_arrays(gd["row_offsets"], gd["column_indices"])
        p = CsrProblem(mark_predecessors=mark_pred).from_host_problem(g)
        for src, expected, _ in gd["cases"]:
            got, _ = run(p, strategy, fused, src)
            assert np.array_equal(got, np.array(expected, dtype=np.int32)), (gd["name"], src)
            if mark_pred:
                assert oracle.check_preds(g.row_offsets, g.column_indices, got, p.extract_results()) == -1
        p.destroy()


@pytest.mark.parametrize("variant", range(len(VARIANTS)), ids=VARIANT_IDS)
def test_random_multigraphs(gpu, variant):
    strategy, fused = VARIANTS[variant]
    rng = np.random.default_rng(1234 + variant)
    for trial in range(6):
        n = int(rng.integers(2, 3000))
        m = int(rng.integers(0, 6 * n))
        g = CsrGraph.from_coo(n, rng.integers(0, n, m), rng.integers(0, n, m),
                              symmetrize=bool(trial % 2), dedup=bool(trial % 3 == 0))
        p = CsrProblem(mark_predecessors=bool(trial % 2)).from_host_problem(g)
        for src in rng.integers(0, n, 3):
            assert_parity(g, p, strategy, fused, src)
        p.destroy()


@pytest.mark.parametrize("variant", range(len(VARIANTS)), ids=VARIANT_IDS)
@pytest.mark.parametrize("cull", [b40c.CULL_EXACT, b40c.CULL_BEST_EFFORT], ids=["exact", "best_effort"])
def test_rmat_scale14(gpu, variant, cull):
    strategy, fused = VARIANTS[variant]
    g = CsrGraph.rmat(14, 16, seed=0xAB)
    p = CsrProblem(mark_predecessors=(variant % 2 == 0)).from_host_problem(g)
    p.set_config(cull_mode=cull)
    # The best-effort discipline lets one duplicate per XCD through a level
    # (plain stores stay in that XCD's write-back L2 until the level
    # barrier), so edge-frontier strategies need more queue headroom.
    qs = 2.0 if cull == b40c.CULL_EXACT else 64.0
    for src in g.sample_sources(3, seed=variant):
        assert_parity(g, p, strategy, fused, src, queue_sizing=qs)
    p.destroy()


@pytest.mark.parametrize("variant", range(len(VARIANTS)), ids=VARIANT_IDS)
def test_gather_tiers_and_grid_shapes(gpu, variant):
    """Small tier thresholds push short rows through the CTA and wave tiers;
    tiny grids and a disabled fast path exercise the barrier loop."""
    strategy, fused = VARIANTS[variant]
    g = CsrGraph.rmat(12, 8, seed=3)
    p = CsrProblem().from_host_problem(g)
    src = int(g.sample_sources(1, seed=9)[0])
    for wave_t, cta_t, grid_t, fast_t, grid in [(1, 2, 3, -1, 0), (2, 5, 40, 64, 3), (64, 65, 100, -1, 1),
                                                (1000000, 1000000, -1, 16, 0), (8, 4096, 2049, 1 << 30, 0),
                                                (64, 1024, 8192, 1024, 0)]:
        # (spread steps off too: small levels then take the tile phases)
        for spread in (1, 2):
            p.set_config(wave_gather_threshold=wave_t, cta_gather_threshold=cta_t,
                         grid_gather_threshold=grid_t, fast_path_threshold=fast_t, spread_steps=spread)
            assert_parity(g, p, strategy, fused, src, max_grid_size=grid)
    p.destroy()


def test_hybrid_switches_back_and_forth(gpu):
    g = CsrGraph.rmat(15, 16, seed=21)
    p = CsrProblem(mark_predecessors=True).from_host_problem(g)
    for sat in [1, 50, 2000, -1]:
        p.set_config(saturation_threshold=sat)
        for src in g.sample_sources(2, seed=sat & 0xFF):
            st = assert_parity(g, p, b40c.HYBRID, True, src)
            if sat == 1:
                assert st.kernel_launches > 1
    p.destroy()


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "iter"])
def test_direction_switch_points(gpu, fused):
    """alpha = 10^6 with beta = 1 flips direction at every level (each
    top-down level is followed by a bottom-up one), stressing every
    hand-over between the two directions; alpha = 1 stays top-down."""
    graphs = [CsrGraph.rmat(15, 16, seed=31), CsrGraph.grid2d(40), CsrGraph.grid3d(9),
              CsrGraph.erdos_renyi(5000, 20000, seed=2)]
    for g in graphs:
        p = CsrProblem(mark_predecessors=True).from_host_problem(g)
        srcs = [0, g.nodes // 3] if g.degrees().min() > 0 else list(g.sample_sources(2, seed=7))
        for alpha, beta in [(1000000, 1), (14, 24), (1, 24), (1000000, 1000000)]:
            p.set_config(do_alpha=alpha, do_beta=beta)
            for src in srcs:
                assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, fused, src)
        p.destroy()


@pytest.mark.parametrize("cull", [b40c.CULL_EXACT, b40c.CULL_BEST_EFFORT], ids=["exact", "best_effort"])
@pytest.mark.parametrize("mark_pred", [False, True], ids=["depth", "pred"])
def test_top_down_after_bottom_up(gpu, cull, mark_pred):
    # The first top-down level after bottom-up ones reads its frontier from
    # the bottom-up bitmap when it is an ordinary step (more than 1024
    # members, fewer than a marked level's edges), else from a queue built
    # from the bitmap.  beta = 1 turns back after every bottom-up level.
    g = CsrGraph.rmat(15, 16, seed=0x3C)
    p = CsrProblem(mark_predecessors=mark_pred).from_host_problem(g)
    qs = g.edges / g.nodes + 1.0
    for alpha, beta in [(1000000, 1), (2, 4), (8, 48)]:
        for spread in (1, 2):
            p.set_config(cull_mode=cull, do_alpha=alpha, do_beta=beta, spread_steps=spread)
            for src in g.sample_sources(3, seed=11):
                assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, True, src, queue_sizing=qs)
    # tiny grids: each workgroup walks many bitmap tiles and spread batches
    p.set_config(do_alpha=1000000, do_beta=1, spread_steps=1)
    for grid in (1, 3):
        for src in g.sample_sources(2, seed=12):
            assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, True, src, queue_sizing=qs, max_grid_size=grid)
    p.destroy()


def test_grid_lattices_small(gpu):
    for w, builder in [(37, CsrGraph.grid2d), (11, CsrGraph.grid3d)]:
        g = builder(w)
        p = CsrProblem().from_host_problem(g)
        for strategy, fused in VARIANTS:
            assert_parity(g, p, strategy, fused, 0)
            assert_parity(g, p, strategy, fused, g.nodes // 2)
        p.destroy()


def test_overflow_is_reported_and_recoverable(gpu):
    g = CsrGraph.rmat(12, 16, seed=5)
    p = CsrProblem().from_host_problem(g)
    src = int(g.sample_sources(1, seed=2)[0])
    p.set_config(cull_mode=b40c.CULL_BEST_EFFORT)
    for strategy, fused in [(b40c.CONTRACT_EXPAND, True), (b40c.TWO_PHASE, False), (b40c.HYBRID, True)]:
        p.reset(b40c.EDGE_FRONTIERS, 0.001)
        with pytest.raises(b40c.BfsError) as ei:
            p.enact(strategy, src, fused=fused)
        assert ei.value.status == 3  # BFS_ERR_OVERFLOW
    p.set_config(cull_mode=b40c.CULL_EXACT)
    assert_parity(g, p, b40c.HYBRID, True, src)
    p.destroy()


@pytest.mark.parametrize("cull", [b40c.CULL_EXACT, b40c.CULL_BEST_EFFORT], ids=["exact", "best_effort"])
@pytest.mark.parametrize("mark_pred", [False, True], ids=["depth", "pred"])
def test_marked_top_down_levels(gpu, cull, mark_pred):
    # Levels of at least max(2^20, V/2) frontier edges mark discoveries with
    # byte stores and compact them behind a barrier instead of claiming;
    # R-MAT scale 17 (4 M edges) has such levels.  Direction-optimizing
    # search with a huge alpha stays top-down on every level.
    g = CsrGraph.rmat(17, 16, seed=0x5A)
    p = CsrProblem(mark_predecessors=mark_pred).from_host_problem(g)
    p.set_config(cull_mode=cull)
    for src in g.sample_sources(2, seed=9):
        assert_parity(g, p, b40c.EXPAND_CONTRACT, True, src)
        p.set_config(do_alpha=1 << 20)
        assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, True, src)
        p.set_config(do_alpha=8)
        assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, True, src)
    # a marked level that overflows the queue reports it, and the problem recovers
    p.reset(b40c.VERTEX_FRONTIERS, 0.02)
    with pytest.raises(b40c.BfsError) as ei:
        p.enact(b40c.EXPAND_CONTRACT, int(g.sample_sources(1, seed=9)[0]))
    assert ei.value.status == 3  # BFS_ERR_OVERFLOW
    assert_parity(g, p, b40c.EXPAND_CONTRACT, True, g.sample_sources(1, seed=9)[0])
    p.destroy()


def hub_fan_graph(hubs=3, leaves=20000, seed=5):
    """Source 0 -> hubs 1..hubs, each hub -> its own leaves, and a random
    sparse mesh over the leaves: level 1 has a few members with tens of
    thousands of edges (more than one workgroup's batch, fewer than a
    marked level's), the shape of an R-MAT search's first levels."""
    rng = np.random.default_rng(seed)
    n = 1 + hubs + hubs * leaves
    src = [np.zeros(hubs, np.int64)]
    dst = [np.arange(1, hubs + 1)]
    for h in range(hubs):
        lv = 1 + hubs + h * leaves + np.arange(leaves)
        src.append(np.full(leaves, 1 + h))
        dst.append(lv)
    a = rng.integers(1 + hubs, n, 2 * leaves)
    b = rng.integers(1 + hubs, n, 2 * leaves)
    src.append(a)
    dst.append(b)
    return CsrGraph.from_coo(n, np.concatenate(src), np.concatenate(dst), symmetrize=True, dedup=True,
                             drop_self_loops=True)


@pytest.mark.parametrize("cull", [b40c.CULL_EXACT, b40c.CULL_BEST_EFFORT], ids=["exact", "best_effort"])
@pytest.mark.parametrize("mark_pred", [False, True], ids=["depth", "pred"])
def test_spread_levels(gpu, cull, mark_pred):
    # Top-down levels of at most 1024 members whose edges exceed one
    # workgroup's batch spread the edges evenly over the grid (spread_steps
    # 1, the default) instead of tile + grid-tier phases (2).  Both forms,
    # both loops, on the hub fan and on R-MAT (whose first levels have it).
    graphs = [hub_fan_graph(), CsrGraph.rmat(16, 16, seed=0x5B)]
    for g in graphs:
        p = CsrProblem(mark_predecessors=mark_pred).from_host_problem(g)
        # best-effort culling queues duplicates (at most a level's edges):
        # a queue of E + V entries cannot overflow
        qs = g.edges / g.nodes + 1.0
        for spread in (1, 2):
            p.set_config(cull_mode=cull, spread_steps=spread)
            # (source 1 is a hub: the first step of the search spreads)
            for src in [0, 1] if g.nodes < 70000 else g.sample_sources(2, seed=3):
                assert_parity(g, p, b40c.EXPAND_CONTRACT, True, src, queue_sizing=qs)
                assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, True, src, queue_sizing=qs)
                p.set_config(do_alpha=1 << 20)
                assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, True, src, queue_sizing=qs)
                p.set_config(do_alpha=8)
        p.destroy()
    # a spread level that overflows the queue reports it, and the problem recovers
    g = graphs[0]
    p = CsrProblem(mark_predecessors=mark_pred).from_host_problem(g)
    p.set_config(cull_mode=cull)
    p.reset(b40c.VERTEX_FRONTIERS, 0.5)
    with pytest.raises(b40c.BfsError) as ei:
        p.enact(b40c.EXPAND_CONTRACT, 0)
    assert ei.value.status == 3  # BFS_ERR_OVERFLOW
    assert_parity(g, p, b40c.EXPAND_CONTRACT, True, 0)
    p.destroy()


@pytest.mark.parametrize("cull", [b40c.CULL_EXACT, b40c.CULL_BEST_EFFORT], ids=["exact", "best_effort"])
@pytest.mark.parametrize("mark_pred", [False, True], ids=["depth", "pred"])
def test_fast_path_visited_modes(gpu, cull, mark_pred):
    # The single-workgroup fast path claims in an LDS copy of the visited
    # bitmap when the graph fits (fast_visited 1, the default) or with
    # atomics on the bitmap in device memory (2).  Bitmap lengths that are
    # not a multiple of four words take the copy's tail loop.
    # Contract-expand, two-phase and the hybrid enter the path at a
    # contraction and hand it back through an expand step; R-MAT's growing
    # frontier makes them leave it mid-search.
    graphs = [CsrGraph.grid2d(61), CsrGraph.erdos_renyi(4133, 20000, seed=0xF1),
              CsrGraph.rmat(14, 8, seed=0x77)]
    for g in graphs:
        p = CsrProblem(mark_predecessors=mark_pred).from_host_problem(g)
        qs = g.edges / g.nodes + 1.0
        # (best-effort grid steps let duplicates through: see test_rmat_scale14)
        qs_edges = 2.0 if cull == b40c.CULL_EXACT else 64.0
        for mode in (1, 2):
            p.set_config(cull_mode=cull, fast_visited=mode)
            for src in g.sample_sources(2, seed=9):
                assert_parity(g, p, b40c.EXPAND_CONTRACT, True, src, queue_sizing=qs)
                assert_parity(g, p, b40c.DIRECTION_OPTIMIZING, True, src, queue_sizing=qs)
                for strategy in (b40c.CONTRACT_EXPAND, b40c.TWO_PHASE, b40c.HYBRID):
                    assert_parity(g, p, strategy, True, src, queue_sizing=qs_edges)
        p.destroy()


@pytest.mark.slow
def test_fast_path_visited_size_limit(gpu):
    # the largest lattice whose visited bitmap fits the fast path's LDS copy
    # and the first that does not; labels by the closed form d = i + j
    assert 1050 * 1050 <= b40c._native.FAST_VISITED_MAX < 1051 * 1051
    for side in (1050, 1051):
        g = CsrGraph.grid2d(side)
        j, k = np.divmod(np.arange(g.nodes), side)
        expect = (j + k).astype(np.int32)
        p = CsrProblem().from_host_problem(g)
        for strategy in (b40c.EXPAND_CONTRACT, b40c.DIRECTION_OPTIMIZING, b40c.CONTRACT_EXPAND,
                         b40c.TWO_PHASE, b40c.HYBRID):
            got, st = run(p, strategy, True, 0)
            assert oracle.compare(got, expect) == -1, (side, strategy)
            assert st.search_depth == 2 * side - 1
        p.destroy()


def test_invalid_source_and_frontier_type(gpu):
    g = CsrGraph.grid2d(4)
    p = CsrProblem().from_host_problem(g)
    p.reset(b40c.EDGE_FRONTIERS, 1.3)
    with pytest.raises(b40c.BfsError):
        p.enact(b40c.HYBRID, 16)
    with pytest.raises(b40c.BfsError):
        p.enact(b40c.EXPAND_CONTRACT, 0)  # needs vertex frontiers
    p.destroy()


def test_repeated_enacts_without_reset(gpu):
    g = CsrGraph.rmat(13, 8, seed=77)
    p = CsrProblem().from_host_problem(g)
    p.reset(b40c.EDGE_FRONTIERS, 1.3)
    row, col = g.row_offsets, g.column_indices
    for src in g.sample_sources(4, seed=4):
        p.enact(b40c.HYBRID, int(src))
        expect, _, _ = oracle.bfs(row, col, int(src))
        assert oracle.compare(p.extract_depths(), expect) == -1
    p.destroy()


# ---------------------------------------------------------------------------
# Full-size configurations
# ---------------------------------------------------------------------------
@pytest.mark.slow
def test_config1_grid2d_1000(gpu):
    g = CsrGraph.grid2d(1000)
    assert g.nodes == 1_000_000 and g.edges == 3_996_000
    j, k = np.divmod(np.arange(g.nodes), 1000)
    expect = (j + k).astype(np.int32)
    p = CsrProblem().from_host_problem(g)
    for strategy, fused in VARIANTS:
        got, st = run(p, strategy, fused, 0)
        assert oracle.compare(got, expect) == -1, (strategy, fused)
        assert st.search_depth == 1999
    p.destroy()


@pytest.mark.slow
def test_config2_rmat22(gpu):
    # every variant on 8 of the 64 bench sources (one oracle run per source)
    g = CsrGraph.rmat(22, 16)
    row, col = g.row_offsets, g.column_indices
    p = CsrProblem().from_host_problem(g)
    srcs = g.sample_sources(64)
    for src in srcs[::8]:
        expect, levels, _ = oracle.bfs(row, col, int(src))
        for strategy, fused in VARIANTS:
            got, st = run(p, strategy, fused, src)
            bad = oracle.compare(got, expect)
            assert bad == -1, (strategy, fused, int(src), bad)
            assert st.search_depth == levels
    p.destroy()


@pytest.mark.slow
def test_config2_rmat22_predecessors(gpu):
    g = CsrGraph.rmat(22, 16)
    row, col = g.row_offsets, g.column_indices
    p = CsrProblem(mark_predecessors=True).from_host_problem(g)
    for src in g.sample_sources(64)[1::16]:
        expect, _, _ = oracle.bfs(row, col, int(src))
        for strategy, fused in [(b40c.EXPAND_CONTRACT, True), (b40c.DIRECTION_OPTIMIZING, True),
                                (b40c.HYBRID, True), (b40c.CONTRACT_EXPAND, True)]:
            got, _ = run(p, strategy, fused, src)
            assert oracle.compare(got, expect) == -1, (strategy, int(src))
            assert oracle.check_preds(row, col, expect, p.extract_results()) == -1, (strategy, int(src))
    p.destroy()


@pytest.mark.slow
def test_config3_grid3d_300(gpu):
    w = 300
    g = CsrGraph.grid3d(w)
    assert g.nodes == 27_000_000 and g.edges == 161_460_000
    idx = np.arange(g.nodes, dtype=np.int64)
    i, rem = np.divmod(idx, w * w)
    jj, kk = np.divmod(rem, w)
    expect = (i + jj + kk).astype(np.int32)
    del idx, i, rem, jj, kk
    p = CsrProblem().from_host_problem(g)
    for strategy, fused in VARIANTS:
        got, st = run(p, strategy, fused, 0)
        assert oracle.compare(got, expect) == -1, (strategy, fused)
        assert st.search_depth == 898
    p.destroy()
