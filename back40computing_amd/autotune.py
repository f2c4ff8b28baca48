"""Autotuner for the BFS tunables (bfs_config).

The reference fixes its tuning at compile time: one KernelPolicy per enactor
for SM2.0 (b40c/graph/bfs/enactor_contract_expand.cuh:351-371,
enactor_expand_contract.cuh:344-364, enactor_two_phase.cuh:636-690,
enactor_hybrid.cuh:571-594).  Here every such choice is a runtime field of
bfs_config, so the tuner measures candidates on the device and keeps the
fastest one whose output is identical to the starting configuration's:

  * tile shape: workgroups per CU, tiles per counter grab, bottom-up tile
    size and grab layout (batched);
  * gather tiers: wave / workgroup / grid row-length thresholds
    (ExpandByWarp / ExpandByCta, two_phase/expand_atomic/cta.cuh:215-393),
    and whether few-member top-down levels spread their edges evenly over
    the grid instead;
  * direction switch points: alpha (top-down -> bottom-up) and beta
    (bottom-up -> top-down) of the direction-optimizing search.

Search: coordinate descent from the library defaults.  Each pass walks the
parameters in order, times every value of one parameter with the others
fixed, and keeps the best; passes repeat until one changes nothing.  A
candidate counts only if its depth labels hash the same as the baseline's
(depths are unique, so any difference is a defect, never a tuning trade).

Results go to a JSON table keyed by device name and workload class; bench.py
and the serving loop load it through tuned_config().
"""
import ctypes
import json
import os
import time
import zlib

from . import _native

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")

# Values tried per parameter, in the order the descent visits them.
SPACE_SINGLE = {
    "blocks_per_cu": [1, 2],
    "tile_grab_max": [1, 4, 16, 32, 64],
    "wave_gather_threshold": [32, 64, 128],
    # rows shorter than a 512-thread workgroup waste most of it in the CTA tier
    "cta_gather_threshold": [256, 512, 1024],
    "grid_gather_threshold": [2048, 4096, 8192, 32768, -1],
    "do_alpha": [2, 4, 8, 14, 24],
    "do_beta": [8, 16, 24, 48],
    "spread_steps": [1, 2],
}

# High-diameter lattices (one search of hundreds to thousands of levels):
# the single-workgroup fast path, the grid shape and the tile grabs matter,
# the direction switch points do not (the search stays top-down).
SPACE_LATTICE = {
    "blocks_per_cu": [1, 2],
    "fast_path_threshold": [256, 512, 1024],
    "tile_grab_max": [1, 4, 16, 32],
    "wave_gather_threshold": [8, 32, 64],
    "cta_gather_threshold": [256, 512, 1024],
    "compaction": [1, 2],
    "fast_visited": [1, 2],
}

SPACE_BATCH = {
    "blocks_per_cu": [1, 2, 4],
    "tile_grab_max": [1, 8, 32, 128],
    # an upper bound: each level shrinks it to give every workgroup tiles
    "bu_tile_blocks": [16, 64, 256],
    "grab_strided": [1, 2],
    "wave_gather_threshold": [32, 64, 128],
    "cta_gather_threshold": [256, 512, 1024],
    "grid_gather_threshold": [256, 512, 1024, 4096],
    "do_alpha": [4, 8, 12],
    "do_beta": [8, 24, 48],
}


def default_config(kind="single"):
    """Library defaults: bfs_config_default, or bfs_batch_config_default."""
    cfg = _native.BfsConfig()
    lib = _native.load()
    (lib.bfs_batch_config_default if kind == "batch" else lib.bfs_config_default)(ctypes.byref(cfg))
    return {f: getattr(cfg, f) for f, _ in _native.BfsConfig._fields_}


def device_name(device=0):
    """Device name from torch when importable, else the gfx target."""
    try:
        import torch
        return torch.cuda.get_device_name(device)
    except Exception:
        return "gfx950"


def _median(xs):
    s = sorted(xs)
    return s[len(s) // 2]


def coordinate_descent(measure, start, space, passes=3, log=None):
    """Minimise measure(config) -> (ms, signature) over `space`.

    measure returns the time of one configuration and a signature of its
    output; candidates whose signature differs from the start's are
    rejected.  Returns (best config, best ms, trial list)."""
    best = dict(start)
    best_ms, ref_sig = measure(best)
    trials = [(dict(best), best_ms, True)]
    if log:
        log(f"start {best_ms:.3f} ms")
    for p in range(passes):
        changed = False
        for name, values in space.items():
            for v in values:
                if best.get(name) == v:
                    continue
                cand = dict(best)
                cand[name] = v
                ms, sig = measure(cand)
                ok = sig == ref_sig
                trials.append((cand, ms, ok))
                if log:
                    log(f"pass {p} {name}={v}: {ms:.3f} ms{'' if ok else ' OUTPUT DIFFERS (rejected)'}")
                if ok and ms < best_ms * 0.99:
                    best, best_ms, changed = cand, ms, True
        if not changed:
            break
    return best, best_ms, trials


def tune_single(problem, sources, strategy=4, reps=3, space=None, passes=3, log=None):
    """Tunes a CsrProblem (device-resident graph) over `sources`.

    Time = median over reps of the summed traversal time of the sources;
    signature = CRC of the depth labels of every source."""
    space = space or SPACE_SINGLE

    def measure(cfg):
        problem.set_config(**cfg)
        sig = 0
        runs = []
        for r in range(reps):
            tot = 0.0
            for s in sources:
                problem.reset(0, 1.0)
                st = problem.enact(strategy, int(s))
                tot += st.elapsed_ms
                if r == 0:
                    sig = zlib.crc32(problem.extract_depths().tobytes(), sig)
            runs.append(tot)
        return _median(runs), sig

    return coordinate_descent(measure, default_config("single"), space, passes, log)


def tune_lattice(problems, strategy=1, reps=3, space=None, passes=3, log=None):
    """Tunes one configuration for several lattice problems (source 0 each).

    Time = median over reps of the summed traversal time; signature = CRC
    of every problem's depth labels."""
    space = space or SPACE_LATTICE

    def measure(cfg):
        sig = 0
        runs = []
        for r in range(reps):
            tot = 0.0
            for p in problems:
                p.set_config(**cfg)
                p.reset(1, 1.3)
                tot += p.enact(strategy, 0).elapsed_ms
                if r == 0:
                    sig = zlib.crc32(p.extract_depths().tobytes(), sig)
            runs.append(tot)
        return _median(runs), sig

    return coordinate_descent(measure, default_config("single"), space, passes, log)


def tune_batch(batch, graph_ids, sources, mode=0, reps=3, space=None, passes=3, check=None, log=None):
    """Tunes a Batch over one set of queries.

    Time = median over reps of the batch's enact time; signature = CRC of
    the depth labels of the queries in `check` (default: every 16th)."""
    space = space or SPACE_BATCH
    check = list(range(0, len(graph_ids), 16)) if check is None else check

    def measure(cfg):
        batch.set_config(**cfg)
        runs = []
        for r in range(reps):
            st = batch.enact(graph_ids, sources, True, mode)
            runs.append(st["elapsed_ms"])
        sig = 0
        for q in check:
            sig = zlib.crc32(batch.depths(q).tobytes(), sig)
        return _median(runs), sig

    return coordinate_descent(measure, default_config("batch"), space, passes, log)


def load_table(path=TABLE_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_entry(kind, cfg, ms, workload, path=TABLE_PATH, device=None):
    table = load_table(path)
    dev = device or device_name()
    table.setdefault(dev, {})[kind] = {
        "config": cfg,
        "ms": round(ms, 4),
        "workload": workload,
        "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime()),
    }
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")


def tuned_config(kind, path=TABLE_PATH, device=None):
    """The tuned bfs_config fields for `kind` ("single", "lattice" or "batch") on this
    device, or {} when the table has none (library defaults apply)."""
    table = load_table(path)
    dev = device or device_name()
    entry = table.get(dev, {}).get(kind)
    if entry is None:
        # any entry for the same gfx target beats none
        for d, kinds in table.items():
            if kind in kinds:
                entry = kinds[kind]
                break
    return dict(entry["config"]) if entry else {}


def main(argv=None):
    import argparse
    import sys

    ap = argparse.ArgumentParser(description="tune bfs_config on this device")
    ap.add_argument("--kind", choices=["single", "lattice", "batch"], required=True)
    ap.add_argument("--scale", type=int, default=22, help="single: R-MAT scale")
    ap.add_argument("--edge-factor", type=int, default=16, help="single: R-MAT edge factor")
    ap.add_argument("--sources", type=int, default=8, help="single: sources per measurement")
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=TABLE_PATH)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    if a.kind == "single":
        from .bfs import CsrGraph, CsrProblem
        g = CsrGraph.rmat(a.scale, a.edge_factor)
        p = CsrProblem().from_host_problem(g)
        p.reset(0, 1.0)
        srcs = g.sample_sources(a.sources)
        best, ms, trials = tune_single(p, srcs, reps=a.reps, passes=a.passes, log=log)
        workload = f"rmat{a.scale}_ef{a.edge_factor}_{a.sources}src direction-optimizing"
        p.destroy()
    elif a.kind == "lattice":
        from .bfs import CsrGraph, CsrProblem
        graphs = [CsrGraph.grid2d(1000), CsrGraph.grid3d(300)]
        problems = [CsrProblem().from_host_problem(g) for g in graphs]
        best, ms, trials = tune_lattice(problems, reps=a.reps, passes=a.passes, log=log)
        workload = "grid2d 1000^2 + grid3d 300^3 from vertex 0, contract-expand"
        for p in problems:
            p.destroy()
    else:
        from .batch import Batch, GraphPool
        from .workloads import batched_rmat
        specs, graphs, gids, srcs = batched_rmat()
        pool = GraphPool()
        for g in graphs:
            pool.add(g)
        b = Batch(pool, max_queries=len(gids))
        best, ms, trials = tune_batch(b, gids, srcs, reps=a.reps, passes=a.passes, log=log)
        workload = "256 queries over 64 R-MAT graphs (scale 14..20, edge factor 8..32), persistent"
        b.destroy()
        pool.destroy()
    save_entry(a.kind, best, ms, workload, path=a.out)
    print(json.dumps({"kind": a.kind, "ms": round(ms, 4), "config": best, "trials": len(trials)}))


if __name__ == "__main__":
    main()
