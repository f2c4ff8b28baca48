"""MI355X (gfx950) breadth-first search: HIP kernels behind a C ABI (libb40c_hip.so).

See back40computing_amd.bfs for the Python mirror of the reference interface.
"""
from . import _native  # noqa: F401
from .bfs import (  # noqa: F401
    CONTRACT_EXPAND, CULL_BEST_EFFORT, CULL_EXACT, DIRECTION_OPTIMIZING, EDGE_FRONTIERS, EXPAND_CONTRACT, HYBRID,
    MIXED_FRONTIERS, TWO_PHASE, VERTEX_FRONTIERS, BfsError, CsrGraph, CsrProblem,
    EnactorContractExpand, EnactorDirectionOptimizing, EnactorExpandContract, EnactorHybrid, EnactorTwoPhase,
    Stats,
    barrier_benchmark, device_count, read_coo, selftest_wave_primitives,
)
