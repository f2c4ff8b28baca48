"""Seeded synthetic workloads of the batched and serving configurations.

batched_rmat: a pool of undirected R-MAT graphs (a=0.57 b=0.19 c=0.19
d=0.05) with scale uniform in [14, 20] and edge factor uniform in [8, 32],
and queries with one random non-isolated source each.

serving_pool: Erdos-Renyi G(n, m = 4n) graphs for n = 2^16 .. 2^23 (average
degree 8) plus one R-MAT graph (scale 24, edge factor 16 at full size), and
Poisson arrival times for a stated rate.
"""
import numpy as np

from .bfs import CsrGraph


def batched_rmat(num_graphs=64, num_queries=256, scale_range=(14, 20), ef_range=(8, 32), seed=0xB47C4):
    rng = np.random.default_rng(seed)
    specs = [(int(rng.integers(scale_range[0], scale_range[1] + 1)), int(rng.integers(ef_range[0], ef_range[1] + 1)))
             for _ in range(num_graphs)]
    graphs = [CsrGraph.rmat(s, ef, seed=seed + 1 + i) for i, (s, ef) in enumerate(specs)]
    gids = rng.integers(0, num_graphs, num_queries).astype(np.int32)
    srcs = np.empty(num_queries, dtype=np.int32)
    for q, gid in enumerate(gids):
        srcs[q] = graphs[gid].sample_sources(1, seed=int(seed * 131 + q))[0]
    return specs, graphs, gids, srcs


SERVING_SEED = 0x5E4E


def serving_pool(er_log_sizes=range(16, 24), rmat_scale=24, rmat_ef=16, seed=SERVING_SEED):
    """Returns (names, graphs)."""
    names, graphs = [], []
    for k in er_log_sizes:
        n = 1 << k
        graphs.append(CsrGraph.erdos_renyi(n, 4 * n, seed=seed + k))
        names.append("er_2^%d" % k)
    if rmat_scale:
        graphs.append(CsrGraph.rmat(rmat_scale, rmat_ef, seed=seed + 1000))
        names.append("rmat%d_ef%d" % (rmat_scale, rmat_ef))
    return names, graphs


def poisson_arrivals(rate_per_s, count, seed=0xA771):
    """Arrival times (seconds from 0) of a Poisson process."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, count))
