#!/usr/bin/env python3
"""BFS throughput on one MI355X: R-MAT scale 22, edge factor 16, 64 sources.

A step is one single-source search (source k of the 64 seeded sources, in
turn) over the device-resident graph: label reset + traversal through the
C ABI of libb40c_hip.so.  value = total traversed edges (E_vis, counted as
test/bfs/test_bfs.cu:360-367) / wall time of the K timed steps.

The roofline block is for the persistent traversal kernel: algorithmic
bytes per launch, B_alg = 8 * E_vis + 20 * V_vis, over the average launch
time measured with HIP events on the library's stream (the enact's
elapsed_ms).  HBM traffic per launch comes from the committed counter
summary under profiles/ when present.

cpu_baseline times the serial reference BFS restated in oracle/ (one
thread, traversal loop only) on a sample of the same sources, and checks the
GPU labels of those sources against it.

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BFS GTEPS (R-MAT s22 ef16) + batched queries/s & p99 vs single-source"
HBM_PEAK_GBS = 8000.0
STRATEGIES = {"ec": 0, "ce": 1, "tp": 2, "hybrid": 3, "do": 4}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--strategy", choices=sorted(STRATEGIES), default="ec")
    ap.add_argument("--iterative", action="store_true", help="per-level launches instead of the persistent kernel")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=12, help="sources timed on the CPU oracle (0: skip)")
    ap.add_argument("--config", default="", help="comma list of bfs_config overrides, e.g. blocks_per_cu=2")
    ap.add_argument("--calibrate-bytes", type=int, default=0,
                    help="also run the plain copy kernel over this many bytes (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this BFS build runs on a single MI355X "
              "(see DESIGN.md for how the search would shard); got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import back40computing_amd as b40c

    if b40c.device_count() < 1:
        raise SystemExit("no HIP device visible")

    t = time.time()
    g = b40c.CsrGraph.rmat(args.scale, args.edge_factor)
    log("graph: R-MAT scale %d ef %d: V=%d E=%d (%.1f s)" % (args.scale, args.edge_factor, g.nodes, g.edges,
                                                             time.time() - t))
    deg = g.degrees()
    srcs = g.sample_sources(args.sources)
    strategy = STRATEGIES[args.strategy]
    fused = not args.iterative
    ft = b40c.bfs._native.load().bfs_strategy_frontier_type(strategy)

    p = b40c.CsrProblem().from_host_problem(g)
    overrides = {}
    for kv in filter(None, args.config.split(",")):
        k, v = kv.split("=")
        overrides[k.strip()] = int(v)
    if overrides:
        p.set_config(**overrides)
    p.reset(ft, 1.3)

    # Untimed pass: traversed-edge counts per source and the labels kept for
    # the CPU parity check.
    e_vis = np.zeros(len(srcs), dtype=np.int64)
    v_vis = np.zeros(len(srcs), dtype=np.int64)
    kept = {}
    for i, s in enumerate(srcs):
        p.enact(strategy, int(s), fused=fused)
        d = p.extract_depths()
        reached = d >= 0
        e_vis[i] = int(deg[reached].sum())
        v_vis[i] = int(reached.sum())
        if i < args.cpu_sample:
            kept[i] = d
    log("per-source E_vis: min %d max %d" % (e_vis.min(), e_vis.max()))

    for k in range(args.warmup):
        p.enact(strategy, int(srcs[k % len(srcs)]), fused=fused)

    b40c.bfs._native.load().bfs_device_synchronize(0)
    kernel_ms = np.zeros(args.steps)
    edges = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        i = k % len(srcs)
        st = p.enact(strategy, int(srcs[i]), fused=fused)
        kernel_ms[k] = st.elapsed_ms
        edges += int(e_vis[i])
    b40c.bfs._native.load().bfs_device_synchronize(0)
    wall = time.perf_counter() - t0

    idx = np.arange(args.steps) % len(srcs)
    gteps = edges / wall / 1e9
    per_src = e_vis[idx] / (kernel_ms * 1e-3) / 1e9
    b_alg = 8.0 * e_vis[idx] + 20.0 * v_vis[idx]
    achieved = float(b_alg.mean() / (kernel_ms.mean() * 1e-3) / 1e9)
    workload = "rmat%d_ef%d_%dsrc" % (args.scale, args.edge_factor, len(srcs))
    traffic = load_traffic(workload)
    traffic_bytes = None
    if traffic and traffic.get("strategy") == args.strategy and traffic.get("fused", True) == fused:
        traffic_bytes = traffic.get("hbm_bytes_per_launch")

    cpu = None
    if args.cpu_sample > 0:
        import oracle
        row, col = g.row_offsets, g.column_indices
        cpu_edges, cpu_ms, parity = 0, 0.0, True
        n = min(args.cpu_sample, len(srcs))
        for i in range(n):
            labels, _, ms = oracle.bfs(row, col, int(srcs[i]))
            cpu_ms += ms
            cpu_edges += int(e_vis[i])
            if oracle.compare(kept[i], labels) != -1:
                parity = False
        cpu = {"value": round(cpu_edges / (cpu_ms * 1e-3) / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": "%d of the %d sources, serial FIFO BFS (oracle/bfs_oracle.c), traversal loop wall time, "
                         "host %s" % (n, len(srcs), cpu_model()),
               "ms_per_source": round(cpu_ms / n, 2), "gpu_labels_match": parity}
        log("cpu baseline: %.3f GTEPS over %d sources, GPU parity %s" % (cpu["value"], n, parity))

    p.destroy()
    if args.calibrate_bytes:
        import ctypes
        ms = ctypes.c_double()
        rc = b40c.bfs._native.load().bfs_copy_calibration(0, args.calibrate_bytes, ctypes.byref(ms))
        log("calibration copy of %d bytes: rc %d, %.3f ms" % (args.calibrate_bytes, rc, ms.value))
    out = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": workload, "graph": "R-MAT a=0.57 b=0.19 c=0.19 d=0.05, symmetrised, "
                   "self-loops and duplicates removed, scrambled ids", "scale": args.scale,
                   "edge_factor": args.edge_factor, "vertices": g.nodes, "directed_edges": g.edges,
                   "sources": len(srcs), "strategy": args.strategy, "fused": fused, "overrides": overrides},
        "gteps_kernel_mean": round(float(per_src.mean()), 3),
        "gteps_kernel_hmean": round(float(len(per_src) / np.sum(1.0 / per_src)), 3),
        "kernel_ms_mean": round(float(kernel_ms.mean()), 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_bytes,
                     "algorithmic_bytes_per_launch": float(b_alg.mean())},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
