"""Property tests of the host graph builders (include/b40c_graph.h)."""
import numpy as np
import pytest

from back40computing_amd import CsrGraph


def rows_of(g):
    r, c = g.row_offsets, g.column_indices
    return [c[r[v]:r[v + 1]].tolist() for v in range(g.nodes)]


def test_grid2d_neighbour_order_matches_reference():
    # b40c/graph/builder/grid2d.cuh:78-100 order: k-1, k+1, j-1, j+1 (self-loop dropped)
    w = 4
    g = CsrGraph.grid2d(w)
    rows = rows_of(g)
    for j in range(w):
        for k in range(w):
            me = j * w + k
            expect = []
            if k - 1 >= 0: expect.append(me - 1)
            if k + 1 < w: expect.append(me + 1)
            if j - 1 >= 0: expect.append(me - w)
            if j + 1 < w: expect.append(me + w)
            assert rows[me] == expect


def test_grid3d_neighbour_order_and_counts():
    w = 5
    g = CsrGraph.grid3d(w)
    assert g.edges == 6 * w * w * (w - 1)
    rows = rows_of(g)
    p = w * w
    for i in range(w):
        for j in range(w):
            for k in range(w):
                me = i * p + j * w + k
                expect = []
                if k > 0: expect.append(me - 1)
                if k < w - 1: expect.append(me + 1)
                if j > 0: expect.append(me - w)
                if j < w - 1: expect.append(me + w)
                if i > 0: expect.append(me - p)
                if i < w - 1: expect.append(me + p)
                assert rows[me] == expect


def test_config_edge_counts():
    # grid2d 1000^2 and grid3d 300^3 closed forms (no self-loops)
    w = 1000
    assert 4 * w * (w - 1) == 3_996_000
    w = 300
    assert 6 * w * w * (w - 1) == 161_460_000


def _check_simple_symmetric(g):
    r, c = g.row_offsets, g.column_indices
    src = np.repeat(np.arange(g.nodes), np.diff(r))
    assert np.all(c != src), "self-loop"
    assert np.all(np.diff(r) >= 0)
    # sorted, unique rows
    for v in range(min(g.nodes, 2000)):
        row = c[r[v]:r[v + 1]]
        assert np.all(np.diff(row) > 0)
    key = src.astype(np.int64) * g.nodes + c
    rkey = c.astype(np.int64) * g.nodes + src
    assert np.array_equal(np.sort(key), np.sort(rkey)), "not symmetric"


def test_rmat_properties_and_determinism():
    a = CsrGraph.rmat(12, 16, seed=99)
    b = CsrGraph.rmat(12, 16, seed=99)
    c = CsrGraph.rmat(12, 16, seed=100)
    assert a.nodes == 4096
    assert np.array_equal(a.row_offsets, b.row_offsets)
    assert np.array_equal(a.column_indices, b.column_indices)
    assert not np.array_equal(a.column_indices[:1000], c.column_indices[:1000])
    assert a.edges <= 2 * 16 * 4096
    _check_simple_symmetric(a)
    deg = a.degrees()
    # skewed: the max degree is far above the mean
    assert deg.max() > 10 * deg.mean()


def test_rmat_thread_count_independent(monkeypatch):
    monkeypatch.setenv("B40C_HOST_THREADS", "1")
    a = CsrGraph.rmat(13, 8, seed=5)
    monkeypatch.setenv("B40C_HOST_THREADS", "7")
    b = CsrGraph.rmat(13, 8, seed=5)
    assert np.array_equal(a.row_offsets, b.row_offsets)
    assert np.array_equal(a.column_indices, b.column_indices)


def test_rmat_quadrant_probabilities():
    # Unscrambled R-MAT at scale 1: 2 vertices, quadrant frequencies a,b,c,d.
    # With scale 1 the only off-diagonal pairs come from quadrants b and c.
    g = CsrGraph.rmat(1, 1 << 16, a=0.57, b=0.19, c=0.19, seed=1, scramble=False)
    assert g.nodes == 2 and g.edges == 2  # only {0,1} survives dedup, both directions
    g = CsrGraph.rmat(16, 4, seed=3, scramble=False)
    deg = g.degrees()
    # vertex 0 is the hub of an unscrambled R-MAT with a > d
    assert deg[0] == deg.max()


def test_erdos_renyi_exact_edge_count():
    n, m = 1 << 12, 4 << 12
    g = CsrGraph.erdos_renyi(n, m, seed=4)
    assert g.edges == 2 * m
    _check_simple_symmetric(g)
    g2 = CsrGraph.erdos_renyi(n, m, seed=4)
    assert np.array_equal(g.column_indices, g2.column_indices)


def test_random_keeps_loops_and_multi_edges():
    g = CsrGraph.random(8, 200, seed=2)
    assert g.edges == 200
    r, c = g.row_offsets, g.column_indices
    src = np.repeat(np.arange(8), np.diff(r))
    assert np.any(src == c)
    pairs = src * 8 + c
    assert len(np.unique(pairs)) < len(pairs)


def test_from_coo_options_and_validation():
    src = np.array([0, 0, 1, 2, 2, 3], dtype=np.int32)
    dst = np.array([1, 1, 2, 2, 0, 0], dtype=np.int32)
    g = CsrGraph.from_coo(4, src, dst)
    assert rows_of(g) == [[1, 1], [2], [0, 2], [0]]
    g = CsrGraph.from_coo(4, src, dst, symmetrize=True, dedup=True, drop_self_loops=True)
    assert rows_of(g) == [[1, 2, 3], [0, 2], [0, 1], [0]]
    with pytest.raises(Exception):
        CsrGraph.from_coo(3, np.array([0, 5], np.int32), np.array([1, 1], np.int32))


def test_sample_sources_distinct_nonzero_degree():
    g = CsrGraph.rmat(12, 4, seed=8)
    s = g.sample_sources(64, seed=1)
    assert len(set(s.tolist())) == 64
    assert np.all(g.degrees()[s] > 0)
    assert np.array_equal(s, g.sample_sources(64, seed=1))
