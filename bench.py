#!/usr/bin/env python3
"""BFS throughput on one MI355X: R-MAT scale 22, edge factor 16, 64 sources.

A step is one sweep over the 64 seeded sources: for each, a single-source
search over the device-resident graph (label reset + traversal through the
C ABI of libb40c_hip.so).  value = total traversed edges (E_vis, counted as
test/bfs/test_bfs.cu:360-367) / wall time of the K timed steps.  The default
strategy is direction-optimizing; the top-down expand-contract rate over the
same sources is reported beside it.

The roofline block is for the persistent traversal kernel: algorithmic
bytes per launch, B_alg = 8 * E_vis + 20 * V_vis (the top-down per-edge and
per-vertex bytes), over the average launch time measured with HIP events on
the library's stream (the enact's elapsed_ms).  HBM traffic per launch comes
from the committed counter summary under profiles/ when present.

batched: 256 queries over 64 R-MAT graphs (scale 14..20, edge factor
8..32) advanced together per level in one persistent launch, against the
same queries run back to back as single-source searches; queries/s and
per-query completion-time percentiles.

serving (--serving): Poisson arrivals against resident Erdos-Renyi graphs
(n = 2^16..2^23, average degree 8) and one R-MAT graph (scale 24, edge factor
16 by default), at several rates; p50/p99 latency and queries/s.

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


def pct(a, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(a), q)), 4) if len(a) else None


def bench_batched(b40c, np, reps, tuned_batch, tuned_single):
    from back40computing_amd.batch import Batch, GraphPool
    from back40computing_amd.workloads import batched_rmat

    t = time.time()
    specs, graphs, gids, srcs = batched_rmat()
    log("batched pool: %d graphs, %d queries (%.1f s)" % (len(graphs), len(gids), time.time() - t))
    pool = GraphPool()
    for g in graphs:
        pool.add(g)
    out = {"workload": "64 R-MAT graphs scale U{14..20} ef U{8..32}, 256 queries",
           "pool_vertices": int(sum(g.nodes for g in graphs)), "pool_edges": int(sum(g.edges for g in graphs))}
    b = Batch(pool, max_queries=len(gids))
    if tuned_batch:
        b.set_config(**tuned_batch)
    for name, mode in [("persistent", Batch.PERSISTENT), ("graph", Batch.GRAPH)]:
        b.enact(gids, srcs, direction_optimizing=True, mode=mode)  # warm-up
        best = None
        for _ in range(reps):
            st = b.enact(gids, srcs, direction_optimizing=True, mode=mode)
            if best is None or st["elapsed_ms"] < best["elapsed_ms"]:
                best = st
                lat = b.latencies_ms()
        out[name] = {"ms": round(best["elapsed_ms"], 4),
                     "queries_per_s": round(len(gids) / (best["elapsed_ms"] * 1e-3), 1),
                     "latency_p50_ms": pct(lat, 50), "latency_p99_ms": pct(lat, 99),
                     "levels": best["levels"], "launches": best["launches"]}
        log("batched %s: %.3f ms for %d queries" % (name, best["elapsed_ms"], len(gids)))
    b.destroy()

    # Back-to-back single-source searches of the same queries (one resident
    # problem per graph, direction-optimizing persistent kernel).
    problems = []
    for g in graphs:
        p = b40c.CsrProblem().from_host_problem(g)
        if tuned_single:
            p.set_config(**tuned_single)
        p.reset(b40c.VERTEX_FRONTIERS, 1.0)
        problems.append(p)
    for q in range(8):
        problems[gids[q]].enact(b40c.DIRECTION_OPTIMIZING, int(srcs[q]))
    b40c.bfs._native.load().bfs_device_synchronize(0)
    kms = np.zeros(len(gids))
    t0 = time.perf_counter()
    done = np.zeros(len(gids))
    for q in range(len(gids)):
        kms[q] = problems[gids[q]].enact(b40c.DIRECTION_OPTIMIZING, int(srcs[q])).elapsed_ms
        done[q] = (time.perf_counter() - t0) * 1e3
    wall = time.perf_counter() - t0
    for p in problems:
        p.destroy()
    pool.destroy()
    out["single_source_back_to_back"] = {
        "ms": round(wall * 1e3, 3), "queries_per_s": round(len(gids) / wall, 1),
        "latency_p50_ms": pct(done, 50), "latency_p99_ms": pct(done, 99),
        "per_query_kernel_ms_p50": pct(kms, 50), "per_query_kernel_ms_p99": pct(kms, 99)}
    out["speedup_queries_per_s"] = round(out["persistent"]["queries_per_s"] /
                                         out["single_source_back_to_back"]["queries_per_s"], 2)
    log("batched speedup over back-to-back: %.2fx" % out["speedup_queries_per_s"])
    return out, graphs, gids, srcs


def bench_serving(b40c, np, rmat_scale, rates, queries_per_rate):
    from back40computing_amd.batch import GraphPool, Server
    from back40computing_amd.workloads import poisson_arrivals, serving_pool

    t = time.time()
    names, graphs = serving_pool(rmat_scale=rmat_scale)
    log("serving pool: %s (%.1f s)" % (", ".join(names), time.time() - t))
    pool = GraphPool()
    for g in graphs:
        pool.add(g)
    rng = np.random.default_rng(0x5E12)
    src_tables = [g.sample_sources(64, seed=7 + i) for i, g in enumerate(graphs)]
    budget = 16 << 30
    results = []
    srv = Server(pool, max_slots=128, levels_per_tick=2, hbm_budget_bytes=budget)
    # page-locked result buffers, allocated and touched before timing
    warm = [[srv._buffer(gi) for _ in range(12)] for gi in range(len(graphs))]
    for gi, bufs in enumerate(warm):
        for b in bufs:
            b.fill(0)
            srv.recycle(gi, b)
    import queue
    import threading
    for rate in rates:
        arrivals = poisson_arrivals(rate, queries_per_rate, seed=int(rate))
        picks = [(int(rng.integers(0, len(graphs))), int(rng.integers(0, 64))) for _ in range(queries_per_rate)]
        before = srv.stats()
        lat, finish, submits = [], [], []
        handed = queue.Queue()

        def consume():
            # waits in submission order and returns each buffer for reuse
            while True:
                item = handed.get()
                if item is None:
                    return
                gi, tk = item
                d, info = srv.wait(tk)
                srv.recycle(gi, d)
                lat.append(info["finish_ms"] - info["submit_ms"])
                finish.append(info["finish_ms"])
                submits.append(info["submit_ms"])

        th = threading.Thread(target=consume)
        th.start()
        t0 = time.perf_counter()
        for i, (gi, si) in enumerate(picks):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            handed.put((gi, srv.submit(gi, int(src_tables[gi][si]))))
        handed.put(None)
        th.join()
        st = srv.stats()
        span = (max(finish) - min(submits)) * 1e-3
        results.append({"offered_rate_per_s": rate, "queries": queries_per_rate,
                        "queries_per_s": round(queries_per_rate / span, 1),
                        "latency_p50_ms": pct(lat, 50), "latency_p99_ms": pct(lat, 99),
                        "max_resident": st["max_resident"], "rejected": st["rejected"] - before["rejected"],
                        "ticks": st["ticks"] - before["ticks"],
                        "tick_ms": round(st["tick_ms"] - before["tick_ms"], 1),
                        "copy_ms": round(st["copy_ms"] - before["copy_ms"], 1),
                        "span_ms": round(span * 1e3, 1)})
        log("serving rate %g/s: %.1f q/s, p50 %.3f ms, p99 %.3f ms" % (
            rate, results[-1]["queries_per_s"], results[-1]["latency_p50_ms"], results[-1]["latency_p99_ms"]))
    srv.destroy()
    pool.destroy()
    return {"graphs": names, "hbm_budget_bytes": budget, "max_slots": 128, "levels_per_tick": 2, "rates": results}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed sweeps over all sources")
    ap.add_argument("--warmup", type=int, default=2, help="untimed sweeps")
    ap.add_argument("--strategy", choices=sorted(STRATEGIES), default="do")
    ap.add_argument("--iterative", action="store_true", help="per-level launches instead of the persistent kernel")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=12, help="sources timed on the CPU oracle (0: skip)")
    ap.add_argument("--config", default="", help="comma list of bfs_config overrides, e.g. blocks_per_cu=2")
    ap.add_argument("--no-tuned", action="store_true",
                    help="library defaults instead of the autotuned table (back40computing_amd/tuned.json)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched section")
    ap.add_argument("--batched-reps", type=int, default=5)
    ap.add_argument("--serving", action="store_true", help="add the serving sweep")
    ap.add_argument("--serving-rmat-scale", type=int, default=24)
    ap.add_argument("--serving-rates", default="100,200,400,800,1600,3200")
    ap.add_argument("--serving-queries", type=int, default=1000)
    ap.add_argument("--calibrate-bytes", type=int, default=0,
                    help="also run the plain copy kernel over this many bytes (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this BFS library runs a search on a single MI355X "
              "(sharding a search across GPUs is not built); got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import back40computing_amd as b40c

    if b40c.device_count() < 1:
        raise SystemExit("no HIP device visible")
    lib = b40c.bfs._native.load()

    t = time.time()
    g = b40c.CsrGraph.rmat(args.scale, args.edge_factor)
    log("graph: R-MAT scale %d ef %d: V=%d E=%d (%.1f s)" % (args.scale, args.edge_factor, g.nodes, g.edges,
                                                             time.time() - t))
    deg = g.degrees()
    srcs = g.sample_sources(args.sources)
    strategy = STRATEGIES[args.strategy]
    fused = not args.iterative

    p = b40c.CsrProblem().from_host_problem(g)
    from back40computing_amd import autotune
    tuned_single = {} if args.no_tuned else autotune.tuned_config("single")
    tuned_batch = {} if args.no_tuned else autotune.tuned_config("batch")
    overrides = {}
    for kv in filter(None, args.config.split(",")):
        k, v = kv.split("=")
        overrides[k.strip()] = int(v)
    if tuned_single or overrides:
        p.set_config(**{**tuned_single, **overrides})
    p.reset(lib.bfs_strategy_frontier_type(strategy), 1.3)

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

    # A step is one sweep over every sampled source, so the timed region
    # covers all of them whatever --steps is.
    for k in range(args.warmup):
        for s in srcs:
            p.enact(strategy, int(s), fused=fused)

    lib.bfs_device_synchronize(0)
    kernel_ms = np.zeros((args.steps, len(srcs)))
    edges = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        for i, s in enumerate(srcs):
            kernel_ms[k, i] = p.enact(strategy, int(s), fused=fused).elapsed_ms
        edges += int(e_vis.sum())
    lib.bfs_device_synchronize(0)
    wall = time.perf_counter() - t0

    kernel_ms = kernel_ms.ravel()
    idx = np.tile(np.arange(len(srcs)), args.steps)
    gteps = edges / wall / 1e9
    per_src = e_vis[idx] / (kernel_ms * 1e-3) / 1e9
    b_alg = 8.0 * e_vis[idx] + 20.0 * v_vis[idx]
    achieved = float(b_alg.mean() / (kernel_ms.mean() * 1e-3) / 1e9)
    workload = "rmat%d_ef%d_%dsrc" % (args.scale, args.edge_factor, len(srcs))
    traffic = load_traffic(workload)
    traffic_bytes = None
    if traffic and fused and args.strategy in traffic.get("strategies", {}):
        traffic_bytes = traffic["strategies"][args.strategy]["hbm_bytes_per_launch"]

    # Top-down expand-contract over the same sources, for reference.
    top_down = None
    if strategy != b40c.EXPAND_CONTRACT:
        p.reset(b40c.VERTEX_FRONTIERS, 1.3)
        tms = [p.enact(b40c.EXPAND_CONTRACT, int(s)).elapsed_ms for s in srcs[:16]]
        top_down = {"strategy": "ec", "fused": True, "sources": 16,
                    "gteps_kernel_mean": round(float(np.mean(e_vis[:16] / (np.array(tms) * 1e-3) / 1e9)), 3),
                    "kernel_ms_mean": round(float(np.mean(tms)), 4)}

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
    del g
    if args.calibrate_bytes:
        import ctypes
        ms = ctypes.c_double()
        rc = lib.bfs_copy_calibration(0, args.calibrate_bytes, ctypes.byref(ms))
        log("calibration copy of %d bytes: rc %d, %.3f ms" % (args.calibrate_bytes, rc, ms.value))

    batched = None
    if not args.no_batched:
        batched, bgraphs, bgids, bsrcs = bench_batched(b40c, np, args.batched_reps, tuned_batch, tuned_single)
        if cpu is not None:
            import oracle
            sample = range(0, len(bgids), 8)
            ms = sum(oracle.bfs(bgraphs[bgids[q]].row_offsets, bgraphs[bgids[q]].column_indices,
                                int(bsrcs[q]))[2] for q in sample)
            cpu["batched_queries_per_s"] = round(len(sample) / (ms * 1e-3), 1)
            cpu["batched_sample"] = "%d of the 256 batched queries, serial oracle, back to back" % len(sample)
        del bgraphs
    serving = None
    if args.serving:
        serving = bench_serving(b40c, np, args.serving_rmat_scale,
                                [float(r) for r in args.serving_rates.split(",")], args.serving_queries)

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
                   "edge_factor": args.edge_factor, "vertices": int(len(deg)), "directed_edges": int(deg.sum()),
                   "sources": len(srcs), "strategy": args.strategy, "fused": fused, "overrides": overrides,
                   "tuned": bool(tuned_single), "tuned_single": tuned_single, "tuned_batch": tuned_batch},
        "gteps_kernel_mean": round(float(per_src.mean()), 3),
        "gteps_kernel_hmean": round(float(len(per_src) / np.sum(1.0 / per_src)), 3),
        "kernel_ms_mean": round(float(kernel_ms.mean()), 4),
        "top_down": top_down,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_bytes,
                     "algorithmic_bytes_per_launch": float(b_alg.mean())},
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
