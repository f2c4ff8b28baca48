"""Graph file readers (b40c_read_dimacs / _metis / _market) on the host.

Pinned by hand-computed pair lists of small files written here (the grammar,
1-based ids, the undirected doubling and its placement, Matrix Market's
column-first reading), by round trips (a generated graph written in each
format reads back to the same CSR), and by rejection of malformed files."""
import numpy as np
import pytest

import back40computing_amd as b40c
from back40computing_amd import CsrGraph, read_coo


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def pairs(kind, path, undirected=False):
    n, s, d = read_coo(kind, path, undirected)
    return n, list(zip(s.tolist(), d.tolist()))


def test_dimacs_golden(tmp_path):
    f = write(tmp_path, "g.gr", "c tiny\np sp 4 3\nc arcs\na 1 2 7\na 2 3 1\na 4 1 2\n")
    assert pairs("dimacs", f) == (4, [(0, 1), (1, 2), (3, 0)])
    # undirected: each arc's reverse comes first (dimacs.cuh:170-180)
    assert pairs("dimacs", f, True) == (4, [(1, 0), (0, 1), (2, 1), (1, 2), (0, 3), (3, 0)])


def test_metis_golden(tmp_path):
    # vertex 1: 2 3; vertex 2: 1; vertex 3: 1; vertex 4: none (blank line)
    f = write(tmp_path, "g.graph", "% comment\n4 2\n2 3\n1\n1\n\n")
    assert pairs("metis", f) == (4, [(0, 1), (0, 2), (1, 0), (2, 0)])
    # header counting all four entries is accepted too (metis.cuh:158-161)
    f2 = write(tmp_path, "g2.graph", "4 4\n2 3\n1\n1\n")
    assert pairs("metis", f2)[1] == [(0, 1), (0, 2), (1, 0), (2, 0)]


def test_market_golden(tmp_path):
    f = write(tmp_path, "g.mtx", "%%MatrixMarket matrix coordinate pattern general\n% c\n3 3 2\n1 2\n3 1 0.5\n")
    # "i j" is the arc j -> i (market.cuh:128-134)
    assert pairs("market", f) == (3, [(1, 0), (0, 2)])
    assert pairs("market", f, True) == (3, [(1, 0), (0, 1), (0, 2), (2, 0)])


@pytest.mark.parametrize("text", [
    "a 1 2 1\np sp 3 1\n",           # arc before the problem line
    "p sp 3 2\na 1 2 1\n",           # fewer arcs than declared
    "p sp 3 1\na 1 2 1\na 2 3 1\n",  # more arcs than declared
    "p sp 3 1\na 1 4 1\n",           # vertex out of range
    "p sp 3 1\na 0 2 1\n",           # ids are 1-based
    "p sp 3 1\na 1 x 1\n",
])
def test_dimacs_rejects_malformed(tmp_path, text):
    with pytest.raises(b40c.BfsError):
        read_coo("dimacs", write(tmp_path, "bad.gr", text))


@pytest.mark.parametrize("text", ["3 1\n2\n1 3\n", "3 1\n4\n1\n", "3 1 1\n2 5\n1 5\n", "x\n"])
def test_metis_rejects_malformed(tmp_path, text):
    with pytest.raises(b40c.BfsError):
        read_coo("metis", write(tmp_path, "bad.graph", text))


@pytest.mark.parametrize("text", ["3 4 1\n1 2\n", "3 3 2\n1 2\n", "3 3 1\n1 4\n", "3 3 1\n1 2\n2 3\n"])
def test_market_rejects_malformed(tmp_path, text):
    with pytest.raises(b40c.BfsError):
        read_coo("market", write(tmp_path, "bad.mtx", text))


def test_missing_file():
    with pytest.raises(b40c.BfsError):
        read_coo("dimacs", "/nonexistent/graph.gr")


def test_round_trips(tmp_path):
    g = CsrGraph.rmat(11, 8, seed=4)
    row, col = g.row_offsets, g.column_indices
    src = np.repeat(np.arange(g.nodes), np.diff(row)).astype(np.int64)
    dst = col.astype(np.int64)
    n, m = g.nodes, g.edges
    # DIMACS: every directed entry as an arc
    d = tmp_path / "g.gr"
    d.write_text("p sp %d %d\n" % (n, m) + "".join("a %d %d 1\n" % (u + 1, v + 1) for u, v in zip(src, dst)))
    # METIS: row k on line k (the graph is symmetric: m = 2 * undirected edges)
    mt = tmp_path / "g.graph"
    mt.write_text("%d %d\n" % (n, m // 2) + "".join(
        " ".join(str(v + 1) for v in col[row[k]:row[k + 1]]) + "\n" for k in range(n)))
    # Matrix Market: "i j" is read as the arc j -> i
    mm = tmp_path / "g.mtx"
    mm.write_text("%%MatrixMarket matrix coordinate pattern general\n%d %d %d\n" % (n, n, m) +
                  "".join("%d %d\n" % (v + 1, u + 1) for u, v in zip(src, dst)))
    for kind, path in [("dimacs", d), ("metis", mt), ("market", mm)]:
        h = CsrGraph.from_file(kind, str(path))
        assert h.nodes == n and h.edges == m, kind
        assert np.array_equal(h.row_offsets, row), kind
        assert np.array_equal(h.column_indices, col), kind
