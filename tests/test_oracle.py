"""Pins the CPU oracle (oracle/bfs_oracle.c) before anything is compared with it.

The reference ships no golden vectors or fixtures for BFS, so the oracle is
pinned by: hand-computed known answers (tests/golden/tiny_graphs.json),
closed forms on grid lattices, an independent brute-force shortest-path
computation on small random graphs, and BFS invariants.  A set of plausible
mutant oracles (wrong order, wrong increment, transposed graph, ...) is run
against the same pins to show each pin set rejects them.
"""
import json
import os

import numpy as np
import pytest

import oracle
from back40computing_amd import CsrGraph

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tiny_graphs.json")


def golden_graphs():
    with open(GOLDEN) as f:
        return json.load(f)["graphs"]


# ---------------------------------------------------------------------------
# Independent checks used as pins
# ---------------------------------------------------------------------------
def brute_force_depths(row, col, src):
    """Bellman-Ford style relaxation to a fixpoint (no queue, no order)."""
    n = len(row) - 1
    inf = np.iinfo(np.int64).max // 4
    d = np.full(n, inf, dtype=np.int64)
    d[src] = 0
    src_of_edge = np.repeat(np.arange(n), np.diff(row))
    col = np.asarray(col, dtype=np.int64)
    for _ in range(n + 1):
        cand = d[src_of_edge] + 1
        new = d.copy()
        if len(col):
            np.minimum.at(new, col, cand)
        if np.array_equal(new, d):
            break
        d = new
    d[d >= inf] = -1
    return d.astype(np.int32)


def floyd_warshall_depths(row, col, src):
    n = len(row) - 1
    inf = 1 << 20
    dist = np.full((n, n), inf, dtype=np.int64)
    np.fill_diagonal(dist, 0)
    for u in range(n):
        for e in range(row[u], row[u + 1]):
            v = col[e]
            if u != v:
                dist[u, v] = 1
    for k in range(n):
        dist = np.minimum(dist, dist[:, k:k + 1] + dist[k:k + 1, :])
    d = dist[src].copy()
    d[d >= inf] = -1
    return d.astype(np.int32)


def invariants_hold(row, col, src, labels):
    """d(src)=0; every edge u->v with d(u)>=0 has 0 <= d(v) <= d(u)+1;
    every d(v)>0 has an in-neighbour at d(v)-1; unreached vertices have no
    reached in-neighbour."""
    n = len(row) - 1
    if labels[src] != 0:
        return False
    has_parent = np.zeros(n, dtype=bool)
    has_parent[src] = True
    for u in range(n):
        du = labels[u]
        for e in range(row[u], row[u + 1]):
            v = col[e]
            if du >= 0:
                if labels[v] < 0 or labels[v] > du + 1:
                    return False
                if labels[v] == du + 1:
                    has_parent[v] = True
    reached = labels >= 0
    return bool(np.all(has_parent[reached]))


# ---------------------------------------------------------------------------
# Known answers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("g", golden_graphs(), ids=lambda g: g["name"])
def test_oracle_golden_vectors(g):
    row = np.array(g["row_offsets"], dtype=np.int32)
    col = np.array(g["column_indices"], dtype=np.int32)
    for src, expected, ev in g["cases"]:
        labels, depth, _ = oracle.bfs(row, col, src)
        assert oracle.compare(labels, np.array(expected, dtype=np.int32)) == -1, (g["name"], src)
        assert depth == max(expected) + 1
        assert oracle.edges_visited(row, labels)[0] == ev
        assert oracle.edges_visited(row, labels)[1] == sum(1 for x in expected if x >= 0)


@pytest.mark.parametrize("w", [1, 2, 3, 17, 64])
def test_oracle_grid2d_closed_form(w):
    g = CsrGraph.grid2d(w)
    assert g.edges == (4 * w * (w - 1))
    for (a, b) in [(0, 0), (w // 2, w // 3), (w - 1, w - 1)]:
        src = a * w + b
        labels, depth, _ = oracle.bfs(g.row_offsets, g.column_indices, src)
        j, k = np.divmod(np.arange(w * w), w)
        expect = np.abs(j - a) + np.abs(k - b)
        assert np.array_equal(labels, expect.astype(np.int32))
        assert depth == int(expect.max()) + 1
        # connected lattice: every directed entry is traversed
        assert oracle.edges_visited(g.row_offsets, labels) == (g.edges, w * w)


@pytest.mark.parametrize("w", [1, 2, 5, 24])
def test_oracle_grid3d_closed_form(w):
    g = CsrGraph.grid3d(w)
    assert g.edges == 6 * w * w * (w - 1)
    for (a, b, c) in [(0, 0, 0), (w // 2, w - 1, w // 3)]:
        src = (a * w + b) * w + c
        labels, _, _ = oracle.bfs(g.row_offsets, g.column_indices, src)
        idx = np.arange(w ** 3)
        i, rem = np.divmod(idx, w * w)
        j, k = np.divmod(rem, w)
        expect = np.abs(i - a) + np.abs(j - b) + np.abs(k - c)
        assert np.array_equal(labels, expect.astype(np.int32))


def test_oracle_level_counts_grid_closed_form():
    # Lattice points of [0,w)^2 with coordinate sum d from the corner.
    w = 40
    g = CsrGraph.grid2d(w)
    labels, _, _ = oracle.bfs(g.row_offsets, g.column_indices, 0)
    counts = oracle.level_counts(labels)
    expect = [min(d, 2 * w - 2 - d) + 1 for d in range(2 * w - 1)]
    assert list(counts) == expect


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    m = int(rng.integers(0, 4 * n + 1))
    src_e = rng.integers(0, n, m)
    dst_e = rng.integers(0, n, m)
    sym = bool(seed % 2)
    g = CsrGraph.from_coo(n, src_e, dst_e, symmetrize=sym, dedup=bool(seed % 3 == 0))
    row, col = g.row_offsets.copy(), g.column_indices.copy()
    s = int(rng.integers(0, n))
    labels, _, _ = oracle.bfs(row, col, s)
    assert np.array_equal(labels, brute_force_depths(row, col, s))
    assert np.array_equal(labels, floyd_warshall_depths(row, col, s))
    assert invariants_hold(row, col, s, labels)


def test_oracle_rmat_invariants():
    g = CsrGraph.rmat(12, 8, seed=7)
    row, col = g.row_offsets.copy(), g.column_indices.copy()
    for s in g.sample_sources(3, seed=3):
        labels, _, _ = oracle.bfs(row, col, int(s))
        assert invariants_hold(row, col, int(s), labels)
        assert np.array_equal(labels, brute_force_depths(row, col, int(s)))


# ---------------------------------------------------------------------------
# Predecessor validation and comparison
# ---------------------------------------------------------------------------
def test_oracle_predecessors_valid_and_validation_rejects_corruptions():
    g = CsrGraph.rmat(10, 8, seed=11)
    row, col = g.row_offsets.copy(), g.column_indices.copy()
    s = int(g.sample_sources(1, seed=5)[0])
    labels, _, _, preds = oracle.bfs(row, col, s, with_preds=True)
    assert oracle.check_preds(row, col, labels, preds) == -1
    assert preds[s] == -2
    reached = np.nonzero(labels > 0)[0]
    v = int(reached[len(reached) // 2])
    # wrong parent depth
    bad = preds.copy()
    bad[v] = v
    assert oracle.check_preds(row, col, labels, bad) == v
    # parent at the right depth but not adjacent
    same_depth_minus1 = np.nonzero(labels == labels[v] - 1)[0]
    non_adj = [int(p) for p in same_depth_minus1 if v not in col[row[p]:row[p + 1]]]
    if non_adj:
        bad = preds.copy()
        bad[v] = non_adj[0]
        assert oracle.check_preds(row, col, labels, bad) == v
    # reached vertex marked unreached / source marked with a parent
    bad = preds.copy()
    bad[v] = -1
    assert oracle.check_preds(row, col, labels, bad) == v
    bad = preds.copy()
    bad[s] = int(col[row[s]])
    assert oracle.check_preds(row, col, labels, bad) == s


def test_oracle_predecessors_of_other_bfs_orders_accepted():
    # Any BFS tree is valid: build the one that takes the LAST eligible parent.
    g = CsrGraph.grid2d(9)
    row, col = g.row_offsets.copy(), g.column_indices.copy()
    labels, _, _ = oracle.bfs(row, col, 40)
    preds = np.full(len(labels), -1, dtype=np.int32)
    preds[40] = -2
    for u in range(len(labels)):
        for e in range(row[u], row[u + 1]):
            v = col[e]
            if labels[v] == labels[u] + 1:
                preds[v] = u
    assert oracle.check_preds(row, col, labels, preds) == -1


def test_oracle_compare_reports_first_mismatch():
    a = np.arange(100, dtype=np.int32)
    assert oracle.compare(a, a.copy()) == -1
    b = a.copy()
    b[37] += 1
    b[80] -= 1
    assert oracle.compare(b, a) == 37


# ---------------------------------------------------------------------------
# Mutant oracles: each plausible mistake must fail at least one pin.
# ---------------------------------------------------------------------------
def _mutant(row, col, src, kind):
    n = len(row) - 1
    lab = np.full(n, -1, dtype=np.int32)
    lab[src] = 0
    if kind == "transposed":
        # walks in-edges instead of out-edges
        rev = [[] for _ in range(n)]
        for u in range(n):
            for e in range(row[u], row[u + 1]):
                rev[col[e]].append(u)
    frontier = [src]
    while frontier:
        u = frontier.pop() if kind == "lifo" else frontier.pop(0)
        nbrs = rev[u] if kind == "transposed" else col[row[u]:row[u + 1]]
        for v in nbrs:
            if kind == "no_visited_check":
                if lab[v] == -1 or lab[v] > lab[u] + 1:
                    pass
                if lab[v] == -1:
                    lab[v] = lab[u] + 1
                    frontier.append(v)
                elif v != src:
                    lab[v] = lab[u] + 1  # overwrites an earlier (smaller) label
            elif lab[v] == -1:
                lab[v] = lab[u] + (2 if kind == "plus_two" else 1)
                frontier.append(v)
    return lab


@pytest.mark.parametrize("kind", ["lifo", "plus_two", "transposed", "no_visited_check"])
def test_pins_reject_mutant_oracles(kind):
    failures = 0
    for g in golden_graphs():
        row = np.array(g["row_offsets"], dtype=np.int32)
        col = np.array(g["column_indices"], dtype=np.int32)
        for src, expected, _ in g["cases"]:
            if not np.array_equal(_mutant(row, col, src, kind), np.array(expected, dtype=np.int32)):
                failures += 1
    grid = CsrGraph.grid2d(6)
    row, col = grid.row_offsets.copy(), grid.column_indices.copy()
    lab = _mutant(row, col, 0, kind)
    j, k = np.divmod(np.arange(36), 6)
    if not np.array_equal(lab, (j + k).astype(np.int32)):
        failures += 1
    if not invariants_hold(row, col, 0, lab):
        failures += 1
    assert failures > 0, kind
