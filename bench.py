#!/usr/bin/env python3
"""BFS throughput on one MI355X: R-MAT scale 22, edge factor 16, 64 sources.

A step is one sweep over the 64 seeded sources: for each, a single-source
search over the device-resident graph (label reset + traversal through the
C ABI of libb40c_hip.so).  value = total traversed edges (E_vis, counted as
test/bfs/test_bfs.cu:360-367) / wall time of the K timed steps.  The default
strategy is direction-optimizing; the top-down expand-contract rate over the
same sources is reported beside it.

roofline: the persistent traversal kernel's algorithmic bytes per launch,
B_alg = 8 * E_vis + 20 * V_vis (the top-down per-edge and per-vertex bytes),
over the average launch time measured with HIP events on the library's
stream (the enact's elapsed_ms).  HBM traffic per launch, counter GB/s and
bytes per traversed edge come from the committed counter summaries under
profiles/ (rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this script), labelled
with the commit they were measured at.

lattices: the 2D 5-point lattice 1000 x 1000 and the 3D 7-point lattice
300^3 from vertex 0 (1999 and 898 levels), every strategy, median of 5
searches each, beside the serial CPU BFS on the same graph.

batched: 256 queries over 64 R-MAT graphs (scale 14..20, edge factor
8..32) advanced together per level in one persistent launch, against the
same queries run back to back as single-source searches; queries/s and
per-query completion-time percentiles, by wall clock and by kernel time.

serving: Poisson arrivals against resident Erdos-Renyi graphs (n = 2^16 ..
2^23, average degree 8) and one R-MAT graph (scale 24, edge factor 16), at
a few stated rates; p50/p99 latency and queries/s.

cpu_baseline times the serial reference BFS restated in oracle/ (one
thread pinned to one core, traversal loop only, median of 5 runs per
source) on a sample of the same sources, and checks the GPU labels of those
sources against it.

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
CPU_RUNS = 5


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


def host_info():
    aff = sorted(os.sched_getaffinity(0))
    return {"cpu": cpu_model(), "nproc": os.cpu_count(), "usable_cores": len(aff)}


class PinnedCore:
    """Pins the calling thread to one core of its allowed set (taskset -c
    <core> for the serial baseline), restoring the set on exit."""

    def __enter__(self):
        self.old = os.sched_getaffinity(0)
        self.core = min(self.old)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.old)


def cpu_bfs_ms(oracle, row, col, src, runs=CPU_RUNS):
    """Median traversal-loop time of `runs` serial searches (pinned), and the labels."""
    import numpy as np
    times, labels = [], None
    with PinnedCore():
        for _ in range(runs):
            labels, _, ms = oracle.bfs(row, col, int(src))
            times.append(ms)
    return float(np.median(times)), labels


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def pct(a, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(a), q)), 4) if len(a) else None


def bench_lattices(b40c, np, oracle, tuned, cpu_runs):
    """Configurations 1 and 3: every strategy from vertex 0, median of 5."""
    out = {}
    for name, builder, width in [("grid2d_1000", b40c.CsrGraph.grid2d, 1000),
                                 ("grid3d_300", b40c.CsrGraph.grid3d, 300)]:
        t = time.time()
        g = builder(width)
        deg = g.degrees()
        p = b40c.CsrProblem().from_host_problem(g)
        if tuned:
            p.set_config(**tuned)
        res = {"vertices": g.nodes, "edges": g.edges, "source": 0, "strategies": {}}
        e_vis = v_vis = None
        for sname, strategy, fused in [("ec", 0, True), ("ce", 1, True), ("tp", 2, True), ("tp_iter", 2, False),
                                       ("hybrid", 3, True), ("do", 4, True)]:
            ft = b40c.VERTEX_FRONTIERS if strategy in (0, 4) else b40c.EDGE_FRONTIERS
            p.reset(ft, 1.3)
            st = p.enact(strategy, 0, fused=fused)  # warm-up
            if e_vis is None:
                d = p.extract_depths()
                e_vis, v_vis = int(deg[d >= 0].sum()), int((d >= 0).sum())
                res["levels"] = st.search_depth
            ms = [p.enact(strategy, 0, fused=fused).elapsed_ms for _ in range(5)]
            med = float(np.median(ms))
            res["strategies"][sname] = {"ms": round(med, 3), "gteps": round(e_vis / (med * 1e-3) / 1e9, 3),
                                        "launches": st.kernel_launches}
        best = min(res["strategies"], key=lambda k: res["strategies"][k]["ms"])
        bms = res["strategies"][best]["ms"]
        b_alg = 8.0 * e_vis + 20.0 * v_vis
        res["best"] = best
        res["roofline"] = {"bound": "hbm", "algorithmic_bytes": b_alg,
                           "achieved": round(b_alg / (bms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(b_alg / (bms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "us_per_level": round(1e3 * bms / res["levels"], 3)}
        if cpu_runs > 0:
            cms, labels = cpu_bfs_ms(oracle, g.row_offsets, g.column_indices, 0, cpu_runs)
            res["cpu_baseline"] = {"ms": round(cms, 3), "gteps": round(e_vis / (cms * 1e-3) / 1e9, 4), "cores": 1,
                                   "runs": cpu_runs, "kind": "port", "pinned": True}
            res["gpu_over_cpu"] = round(cms / bms, 2)
            p.reset(b40c.VERTEX_FRONTIERS, 1.3)
            p.enact(0, 0)
            res["gpu_labels_match"] = oracle.compare(p.extract_depths(), labels) == -1
        p.destroy()
        del g
        out[name] = res
        log("%s: best %s %.3f ms (%.1f us/level)%s (%.1f s)" % (
            name, best, bms, res["roofline"]["us_per_level"],
            ", cpu %.3f ms" % res["cpu_baseline"]["ms"] if "cpu_baseline" in res else "", time.time() - t))
    return out


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
            t0 = time.perf_counter()
            st = b.enact(gids, srcs, direction_optimizing=True, mode=mode)
            call_ms = (time.perf_counter() - t0) * 1e3
            if best is None or st["elapsed_ms"] < best["elapsed_ms"]:
                best = st
                best_call = call_ms
                lat = b.latencies_ms()
        out[name] = {"ms": round(best["elapsed_ms"], 4),
                     "queries_per_s": round(len(gids) / (best["elapsed_ms"] * 1e-3), 1),
                     # the whole call: query seeding, the level loop, the
                     # statistics read-back
                     "call_ms": round(best_call, 4),
                     "call_queries_per_s": round(len(gids) / (best_call * 1e-3), 1),
                     "latency_p50_ms": pct(lat, 50), "latency_p99_ms": pct(lat, 99),
                     "levels": best["levels"], "launches": best["launches"]}
        log("batched %s: %.3f ms for %d queries (call %.3f ms)" % (name, best["elapsed_ms"], len(gids), best_call))
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
        "kernel_ms_sum": round(float(kms.sum()), 3),
        "latency_p50_ms": pct(done, 50), "latency_p99_ms": pct(done, 99),
        "per_query_kernel_ms_p50": pct(kms, 50), "per_query_kernel_ms_p99": pct(kms, 99)}
    pers = out["persistent"]
    # wall-clock ratio (each side's host overheads included) and the
    # kernel-only ratio (batch level loop vs the sum of single-search kernels)
    out["speedup_queries_per_s"] = round(pers["queries_per_s"] / out["single_source_back_to_back"]["queries_per_s"], 2)
    out["speedup_call_vs_wall"] = round(wall * 1e3 / pers["call_ms"], 2)
    out["speedup_kernel_only"] = round(float(kms.sum()) / pers["ms"], 2)
    log("batched speedup over back-to-back: %.2fx wall, %.2fx call-inclusive, %.2fx kernel-only" % (
        out["speedup_queries_per_s"], out["speedup_call_vs_wall"], out["speedup_kernel_only"]))
    return out, graphs, gids, srcs


def bench_serving(b40c, np, oracle, rmat_scale, rates, queries_per_rate, cpu_runs):
    from back40computing_amd.batch import GraphPool, Server
    from back40computing_amd.workloads import poisson_arrivals, serving_pool

    from back40computing_amd.primitives import DeviceCsr
    from back40computing_amd.workloads import SERVING_SEED

    # Erdos-Renyi graphs from the host builder; the R-MAT graph sampled,
    # sorted and deduplicated on the device (the same CSR as the host
    # builder's, pinned by the device-construction tests) and added to the
    # pool without a host round trip; its host copy only feeds the CPU
    # baseline and the source table.
    t = time.time()
    names, graphs = serving_pool(rmat_scale=0)
    build_s = time.time() - t
    t = time.time()
    pool = GraphPool()
    for g in graphs:
        pool.add(g)
    upload_s = time.time() - t
    device_build = None
    if rmat_scale:
        t = time.time()
        dg = DeviceCsr.rmat(rmat_scale, 16, seed=SERVING_SEED + 1000)
        wall = time.time() - t
        pool.add_device(dg)
        device_build = {"graph": "rmat%d_ef16" % rmat_scale, "vertices": dg.nodes, "directed_edges": dg.edges,
                        "device_ms": round(dg.build_ms, 1), "wall_s": round(wall, 2)}
        names.append("rmat%d_ef16" % rmat_scale)
        graphs.append(dg.to_host())
        dg.free()
    log("serving pool: %s (host builds %.1f s, R-MAT built on the device in %s ms)" % (
        ", ".join(names), build_s, device_build["device_ms"] if device_build else "-"))
    rng = np.random.default_rng(0x5E12)
    src_tables = [g.sample_sources(64, seed=7 + i) for i, g in enumerate(graphs)]
    budget = 16 << 30
    max_slots, levels_per_tick = 128, 2
    results = []
    srv = Server(pool, max_slots=max_slots, levels_per_tick=levels_per_tick, hbm_budget_bytes=budget)
    # page-locked result buffers, allocated and touched before timing
    warm = [[srv._buffer(gi) for _ in range(12)] for gi in range(len(graphs))]
    for gi, bufs in enumerate(warm):
        for b in bufs:
            b.fill(0)
            srv.recycle(gi, b)
    import queue
    import threading
    all_picks = []
    for rate in rates:
        arrivals = poisson_arrivals(rate, queries_per_rate, seed=int(rate))
        picks = [(int(rng.integers(0, len(graphs))), int(rng.integers(0, 64))) for _ in range(queries_per_rate)]
        all_picks += picks
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
    out = {"graphs": names, "vertices": [g.nodes for g in graphs], "edges": [g.edges for g in graphs],
           "host_build_s": round(build_s, 2), "upload_s": round(upload_s, 2), "device_build": device_build,
           "hbm_budget_bytes": budget, "max_slots": max_slots, "levels_per_tick": levels_per_tick,
           "rates": results}
    if cpu_runs > 0:
        # serial time of one query per graph (the first pick on it), weighted
        # by how often the offered mix picks each graph
        per_graph = []
        for gi, g in enumerate(graphs):
            ms, _ = cpu_bfs_ms(oracle, g.row_offsets, g.column_indices, int(src_tables[gi][0]),
                               1 if g.edges > 1e8 else cpu_runs)
            per_graph.append(ms)
        freq = np.bincount([gi for gi, _ in all_picks], minlength=len(graphs)) / max(1, len(all_picks))
        mean_ms = float(np.dot(freq, per_graph))
        out["cpu_baseline"] = {"queries_per_s": round(1e3 / mean_ms, 2), "per_graph_ms": [round(x, 2) for x in per_graph],
                               "cores": 1, "kind": "port", "pinned": True,
                               "sample": "one source per graph, median of %d runs (1 run for graphs over 1e8 edges), "
                                         "weighted by the offered mix" % cpu_runs}
    return out


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
    ap.add_argument("--cpu-sample", type=int, default=4, help="sources timed on the CPU oracle (0: skip all CPU legs)")
    ap.add_argument("--cpu-runs", type=int, default=CPU_RUNS, help="serial runs per CPU sample (median)")
    ap.add_argument("--config", default="", help="comma list of bfs_config overrides, e.g. blocks_per_cu=2")
    ap.add_argument("--no-tuned", action="store_true",
                    help="library defaults instead of the autotuned table (back40computing_amd/tuned.json)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched section")
    ap.add_argument("--batched-reps", type=int, default=5)
    ap.add_argument("--no-lattices", action="store_true", help="skip the lattice section")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sweep")
    ap.add_argument("--serving-rmat-scale", type=int, default=24)
    ap.add_argument("--serving-rates", default="200,1000")
    ap.add_argument("--serving-queries", type=int, default=200)
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
    oracle = None
    if args.cpu_sample > 0:
        import oracle

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
    traffic_bytes, measured = None, None
    if traffic and fused and args.strategy in traffic.get("strategies", {}):
        tr = traffic["strategies"][args.strategy]
        traffic_bytes = tr["hbm_bytes_per_launch"]
        measured = {"profiled_at": traffic.get("commit"),
                    "hbm_bytes_per_launch": traffic_bytes,
                    "bytes_per_traversed_edge": round(traffic_bytes / float(e_vis.mean()), 3),
                    # counter bytes over this run's average launch time
                    "counter_gbs": round(traffic_bytes / (kernel_ms.mean() * 1e-3) / 1e9, 2),
                    "counter_frac": round(traffic_bytes / (kernel_ms.mean() * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        for k in ("kernel_us", "waves_per_cu", "grbm_gui_active"):
            if k in tr:
                measured[k] = tr[k]
        if "phases" in traffic:
            # per-level launches of the two-phase strategy: each phase's
            # counter bandwidth (profiles/step_kernels_rmat22_tp_iterative.md)
            measured["phases"] = {"workload": "two-phase per-level launches",
                                  "profiled_at": traffic["phases"].get("commit")}
            for ph, v in traffic["phases"]["phases"].items():
                measured["phases"][ph] = {"avg_us": v["avg_us"], "counter_gbs": v["counter_gbs"],
                                          "counter_frac": v["counter_frac"]}

    # Top-down expand-contract over the same sources, for reference.
    top_down = None
    if strategy != b40c.EXPAND_CONTRACT:
        p.reset(b40c.VERTEX_FRONTIERS, 1.3)
        tms = [p.enact(b40c.EXPAND_CONTRACT, int(s)).elapsed_ms for s in srcs]
        top_down = {"strategy": "ec", "fused": True, "sources": len(srcs),
                    "gteps_kernel_mean": round(float(np.mean(e_vis / (np.array(tms) * 1e-3) / 1e9)), 3),
                    "kernel_ms_mean": round(float(np.mean(tms)), 4)}
        if traffic and "ec" in traffic.get("strategies", {}):
            tb = traffic["strategies"]["ec"]["hbm_bytes_per_launch"]
            top_down["measured"] = {"profiled_at": traffic.get("commit"), "hbm_bytes_per_launch": tb,
                                    "bytes_per_traversed_edge": round(tb / float(e_vis.mean()), 3),
                                    "counter_gbs": round(tb / (np.mean(tms) * 1e-3) / 1e9, 2)}

    cpu = None
    if args.cpu_sample > 0:
        row, col = g.row_offsets, g.column_indices
        cpu_edges, cpu_ms, parity = 0, 0.0, True
        n = min(args.cpu_sample, len(srcs))
        for i in range(n):
            ms, labels = cpu_bfs_ms(oracle, row, col, int(srcs[i]), args.cpu_runs)
            cpu_ms += ms
            cpu_edges += int(e_vis[i])
            if oracle.compare(kept[i], labels) != -1:
                parity = False
        cpu = {"value": round(cpu_edges / (cpu_ms * 1e-3) / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": "%d of the %d sources, serial FIFO BFS (oracle/bfs_oracle.c), traversal loop wall time, "
                         "one thread pinned to one core, median of %d runs per source" % (n, len(srcs), args.cpu_runs),
               "host": host_info(), "ms_per_source": round(cpu_ms / n, 2), "gpu_labels_match": parity}
        log("cpu baseline: %.3f GTEPS over %d sources, GPU parity %s" % (cpu["value"], n, parity))

    p.destroy()
    del g
    if args.calibrate_bytes:
        import ctypes
        ms = ctypes.c_double()
        rc = lib.bfs_copy_calibration(0, args.calibrate_bytes, ctypes.byref(ms))
        log("calibration copy of %d bytes: rc %d, %.3f ms" % (args.calibrate_bytes, rc, ms.value))

    lattices = None
    if not args.no_lattices:
        # the lattice class of the tuned table when it has one
        tuned_lattice = {} if args.no_tuned else (autotune.tuned_config("lattice") or tuned_single)
        lattices = bench_lattices(b40c, np, oracle, tuned_lattice, args.cpu_runs if args.cpu_sample > 0 else 0)

    batched = None
    if not args.no_batched:
        batched, bgraphs, bgids, bsrcs = bench_batched(b40c, np, args.batched_reps, tuned_batch, tuned_single)
        if cpu is not None:
            sample = range(0, len(bgids), 8)
            with PinnedCore():
                ms = sum(oracle.bfs(bgraphs[bgids[q]].row_offsets, bgraphs[bgids[q]].column_indices,
                                    int(bsrcs[q]))[2] for q in sample)
            cpu["batched_queries_per_s"] = round(len(sample) / (ms * 1e-3), 1)
            cpu["batched_sample"] = "%d of the 256 batched queries, serial oracle, back to back, pinned" % len(sample)
        del bgraphs
    serving = None
    if not args.no_serving:
        serving = bench_serving(b40c, np, oracle, args.serving_rmat_scale,
                                [float(r) for r in args.serving_rates.split(",")], args.serving_queries,
                                args.cpu_runs if args.cpu_sample > 0 else 0)

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
                   "sources": len(srcs), "step": "one search from each of the %d sources" % len(srcs),
                   "searches_timed": args.steps * len(srcs), "strategy": args.strategy, "fused": fused,
                   "overrides": overrides, "tuned": bool(tuned_single), "tuned_single": tuned_single,
                   "tuned_batch": tuned_batch},
        "gteps_kernel_mean": round(float(per_src.mean()), 3),
        "gteps_kernel_hmean": round(float(len(per_src) / np.sum(1.0 / per_src)), 3),
        "kernel_ms_mean": round(float(kernel_ms.mean()), 4),
        "top_down": top_down,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_bytes,
                     "algorithmic_bytes_per_launch": float(b_alg.mean()), "measured": measured},
        "lattices": lattices,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
