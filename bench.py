#!/usr/bin/env python3
"""Benchmark of the TWoS estimator hot path on one MI355X.

Workload: the 81,920-triangle seeded spherical-harmonic icosphere, 16,384
interior sites, independent primal (g = x + 0.3 y^2) and adjoint (g = z) Philox
streams, binary32 walks with a binary64 combine, eps 1e-4, adaptive per-site
budget 256 -> 16,384 walks per side until 1.96 * SE <= 0.02 |TD|.

One step = one full estimate over all sites (every site runs its doubling
ladder; all sites share one persistent launch per round, captured as a
hipGraph).  The metric is TD sites/s: sites finished at the CI target per
second of wall time.  Beside it: walk-steps/s, the plain one-launch-per-site
GPU schedule on a seeded 512-site sample, the walk kernel against the HBM
roofline, and the CPU oracle (same streams, OpenMP over all host cores) on a
seeded 320-site sample (about 13 s on 16 threads).

    python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--workload sh_icosphere|cooling_block|unit_ball]
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "walking-into-holes_amd"))

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK = 78.6e12
FP32_VALU_PEAK = 157.3e12
NODE_BYTES = 128  # one 4-wide node: four child boxes + four child links
TRI_BYTES = 36    # nine binary32 coordinates per triangle test
GRID_BYTES = 4    # one binary32 distance per sample of a distance-grid lookup (1 centre or 8 corners)
TUNING_CACHE = ROOT / "walking-into-holes_amd" / "twos" / "tuning_cache.json"


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sh_icosphere",
                    choices=["sh_icosphere", "cooling_block", "unit_ball", "serve", "copy", "design", "diffwos",
                             "certify"])
    ap.add_argument("--design-rounds", type=int, default=2, help="design: insertion rounds")
    ap.add_argument("--design-candidates", type=int, default=65536, help="design: candidate pool size")
    ap.add_argument("--design-scout-walks", type=int, default=4096, help="design: scout walks per side")
    ap.add_argument("--requests", type=int, default=300, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=40.0, help="serve: Poisson arrivals per second")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the short serving stream and design loop the default line carries")
    ap.add_argument("--plain-sites", type=int, default=512)
    ap.add_argument("--cpu-sites", type=int, default=160)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--waves-per-simd", type=int, default=-1,
                    help="persistent occupancy variant; default: the autotuner's choice")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shape instead of the tuner's")
    ap.add_argument("--retune", action="store_true",
                    help="run the autotuner's trials instead of reading its committed cache "
                         "(writes the file named by --tuning-cache)")
    ap.add_argument("--tuning-cache", default="tuning_cache.retuned.json",
                    help="with --retune: the cache file to write (default: in the working directory)")
    ap.add_argument("--block", type=int, default=0, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--lds-stack", type=int, default=0, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--step-batch", type=int, default=-1, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--refill-batch", type=int, default=0, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--grid-kappa", type=float, default=0.0, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--grid-cells", type=int, default=0, help="fixed distance-grid resolution (skips the autotuner)")
    ap.add_argument("--grid-corners", type=int, default=-1,
                    help="fixed distance-grid mode: 0 centres, 1 corners, 2 binary16 corner records")
    ap.add_argument("--exact", action="store_true", help="exact closest-point radius at every mesh step")
    ap.add_argument("--sites", type=int, default=0, help="override the site count (profiling runs)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the live rocprofv3 --pmc traffic passes (roofline.traffic from the committed profile)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="internal: one eager estimate of the workload at its tuned shape, for a counter pass")
    ap.add_argument("--fp64-sites", type=int, default=256, help="binary64 rerun sample (0: skip)")
    return ap.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str):
    from twos import workloads as W
    if name == "sh_icosphere":
        return W.sh_icosphere()
    if name == "cooling_block":
        return W.cooling_block()
    return W.unit_ball()


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        return 2
    import twos
    from twos import workloads as W

    if args.workload == "serve":
        return serve(args)
    if args.workload == "copy":
        return copy_probe()
    if args.traffic_probe:
        return traffic_probe(args)
    if args.workload in ("diffwos", "certify"):
        return f4_bench(args)
    if args.workload == "design":
        return design_bench(args)
    w = build_workload(args.workload)
    sites = w.sites if args.sites <= 0 else w.sites[:args.sites]
    n = len(sites)
    log(f"workload {w.name}: {n} sites")
    ctx = twos.Context()
    if w.container == twos.MESH:
        mesh = ctx.mesh(w.vertices, w.faces)
        geom = ctx.geometry(twos.MESH, mesh=mesh)
        inside = ctx.inside(mesh, sites)
        if not inside.all():
            raise SystemExit(f"{int((~inside).sum())} sites outside the mesh")
        log(f"LBVH built in {mesh.build_ms:.2f} ms, depth {mesh.depth}")
    else:
        geom = ctx.geometry(w.container, ball=w.ball, box_lo=w.box_lo, box_hi=w.box_hi, spheres=w.spheres,
                            cylinders=w.cylinders)
        mesh = None
    sides = W.sides(w, twos.side)
    adaptive = w.walks_max > w.walks
    grid_on = w.container == twos.MESH and not args.exact
    kw = dict(combine=twos.CROSS, precision=w.precision, seed=2026, walks=w.walks, walks_max=w.walks_max,
              ci_rel=w.ci_rel, eps=w.eps, max_steps=w.max_steps, distance_grid=grid_on)
    # Launch shape of the persistent path: the autotuner's choice (untimed,
    # before warmup), unless fixed on the command line.
    shape = {}
    tuning = None
    grid_build_ms = 0.0
    fixed = {k: v for k, v in (("waves_per_simd", args.waves_per_simd), ("block", args.block),
                               ("lds_stack", args.lds_stack), ("grid_kappa", args.grid_kappa),
                               ("step_batch", args.step_batch), ("refill_batch", args.refill_batch))
             if (v >= 0 if k in ("waves_per_simd", "step_batch") else v > 0)}
    if fixed or args.grid_cells > 0 or args.grid_corners >= 0:
        shape = fixed
        if (args.grid_cells > 0 or args.grid_corners >= 0) and mesh is not None:
            grid_build_ms = ctx.set_distance_grid(mesh, args.grid_cells if args.grid_cells > 0 else 128,
                                                  max(args.grid_corners, 0))
    elif not args.no_autotune:
        from twos import autotune
        t0 = time.perf_counter()
        tuning = autotune.autotune(ctx, geom, sites, sides, geom_sig=f"{w.name}", combine=twos.CROSS,
                                   precision=w.precision, eps=w.eps, max_steps=w.max_steps,
                                   distance_grid=grid_on, mesh=mesh, log=log,
                                   cache_path=Path(args.tuning_cache) if args.retune else TUNING_CACHE)
        shape = {k: v for k, v in tuning.choice.items()
                 if k not in ("grid_cells", "grid_corners") and (grid_on or k != "grid_kappa")}
        grid_build_ms = tuning.grid_build_ms
        log(f"autotune {time.perf_counter() - t0:.1f} s: {tuning.choice}, "
            f"{tuning.speedup:.3f}x the default shape per path")

    def run(schedule, pts, graph):
        return ctx.estimate(geom, pts, sides, schedule=schedule, use_graph=graph, **kw, **shape)

    def finished(res):
        if adaptive:
            return int(np.sum((res.status == 0) & (1.96 * res.score_se <= w.ci_rel * np.abs(res.score))))
        return len(res.score)

    graph = not args.no_graph
    for i in range(args.warmup):
        res = run(twos.PERSISTENT, sites, graph)
        log(f"warmup {i}: {res.stats['seconds']:.3f} s, rounds {res.stats['rounds']}")
    times, walk_ms, steps, paths, done = [], [], [], [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = run(twos.PERSISTENT, sites, graph)
        times.append(time.perf_counter() - t0)
        walk_ms.append(res.stats["walk_kernel_ms"])
        steps.append(res.stats["steps"])
        paths.append(res.stats["paths"])
        done.append(finished(res))
        log(f"step {i}: {times[-1]:.3f} s, {done[-1]} sites at target, steps {steps[-1]:.3e}")
    total_t = float(np.sum(times))
    value = float(np.sum(done)) / total_t
    steps_per_s = float(np.sum(steps)) / total_t
    capped = int(np.sum(res.status == 1))

    out_shape = {"block": 256, "waves_per_simd": 0, "lds_stack": 24, "step_batch": 0, "refill_batch": 8,
                         **({"grid_kappa": 4.0, "grid_cells": 128, "grid_corners": 0} if grid_on else {}), **shape,
                         **({k: tuning.choice[k] for k in ("grid_cells", "grid_corners")}
                            if tuning is not None and "grid_cells" in tuning.choice else {}),
                         **({"grid_cells": args.grid_cells} if grid_on and args.grid_cells > 0 else {}),
                         **({"grid_corners": args.grid_corners} if grid_on and args.grid_corners >= 0 else {})}
    # Walk kernel alone (eager launches, HIP events on the library's stream).
    eager = run(twos.PERSISTENT, sites, False)
    eager_ms = eager.stats["walk_kernel_ms"]
    checksum = {"steps": int(np.sum(eager.steps)), "walks": int(np.sum(eager.walks)),
                "score_sha16": hashlib.sha256(np.ascontiguousarray(eager.score).tobytes()).hexdigest()[:16]}
    launches = eager.stats["walk_launches"]
    queries = eager.stats["steps"] + eager.stats["paths"]
    roofline = None
    if w.container == twos.MESH:
        sub = sites[:: max(1, n // 1024)]
        cnt = ctx.estimate(geom, sub, sides, schedule=twos.PERSISTENT, count_traversal=True,
                           **{**kw, **shape, "walks": 512, "walks_max": 0})
        q_sub = cnt.stats["steps"] + cnt.stats["paths"]
        n_nodes = cnt.stats["nodes_visited"] / q_sub
        n_tris = cnt.stats["tris_tested"] / q_sub
        n_grid = cnt.stats["grid_lookups"] / q_sub
        exact = 1.0 - cnt.stats["grid_hits"] / q_sub  # queries that traverse the BVH
        # Samples per grid lookup: one binary32 centre, eight binary32 corners,
        # or one 16-byte record of eight binary16 corners.
        grid_bytes = {0: GRID_BYTES, 1: 8 * GRID_BYTES, 2: 16}[out_shape.get("grid_corners", 0)]
        b_query = NODE_BYTES * n_nodes + TRI_BYTES * n_tris + grid_bytes * n_grid
        achieved = queries * b_query / (eager_ms * 1e-3) / 1e9
        if args.no_traffic:
            traffic, traffic_source = measured_traffic(w.name, n, out_shape, launches), "committed profile"
        else:
            traffic, traffic_source = live_traffic(w.name, n, out_shape, launches)
            if traffic is None:
                committed = measured_traffic(w.name, n, out_shape, launches)
                if committed is not None:
                    traffic, traffic_source = committed, f"committed profile ({traffic_source})"
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "algorithmic_bytes_per_launch": round(queries * b_query / launches),
                    "bytes_per_query": round(b_query, 1), "nodes_per_query": round(n_nodes, 2),
                    "tris_per_query": round(n_tris, 2), "grid_lookups_per_query": round(n_grid, 3),
                    "traversing_query_fraction": round(exact, 3),
                    "nodes_per_traversing_query": round(n_nodes / max(exact, 1e-9), 2),
                    # Visits to the top three levels of the 4-wide tree (21
                    # nodes), the share a per-block LDS copy of them could serve.
                    "top_level_nodes_per_query": round(cnt.stats["top_nodes_visited"] / q_sub, 4),
                    "top_level_node_fraction": round(cnt.stats["top_nodes_visited"]
                                                     / max(cnt.stats["nodes_visited"], 1), 4),
                    "walk_kernel_ms_per_launch": round(eager_ms / launches, 3),
                    "launches": launches, "traffic_source": traffic_source,
                    "traffic_x_algorithmic": (round(traffic / (queries * b_query / launches), 3)
                                              if traffic else None)}
        frozen = frozen_unit(w.name)
        if frozen is not None:
            # The same queries priced at the bytes an exact BVH closest-point
            # query costs (the unit frozen from the oracle's BVH): how much
            # query work per second the walk kernel delivers in the terms of
            # the reference's Embree query, against the same 8 TB/s.
            eq = queries * frozen["frozen_bytes_per_query"] / (eager_ms * 1e-3) / 1e9
            roofline.update({"frozen_bytes_per_query": round(frozen["frozen_bytes_per_query"], 1),
                             "achieved_frozen": round(eq, 1), "frac_frozen": round(eq / HBM_PEAK_GBS, 4),
                             "frozen_unit": "tests/golden/roofline_c3.json (oracle BVH, exact-walk query mix)"})
    else:
        flops = 23 if w.precision == 64 else 177
        peak = FP64_VALU_PEAK if w.precision == 64 else FP32_VALU_PEAK
        achieved = queries * flops / (eager_ms * 1e-3) / 1e12
        roofline = {"bound": "valu", "achieved": round(achieved, 2), "peak": peak / 1e12, "unit": "TFLOP/s",
                    "frac": round(achieved * 1e12 / peak, 4), "traffic": None, "flops_per_query": flops,
                    "walk_kernel_ms_per_launch": round(eager_ms / launches, 3), "launches": launches}

    out = {
        "metric": "TD sites/s at 95% CI half-width <= 2%|TD|",
        "value": round(value, 2),
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total_t / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 walks, f64 combine" if w.precision == 32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {n} sites, eps {w.eps}, CROSS primal/adjoint Philox streams, "
                               + (f"walks {w.walks}->{w.walks_max} per side until 1.96 SE <= {w.ci_rel}|TD|"
                                  if adaptive else f"{w.walks} walks per side"),
                   "sites": n, "schedule": "persistent + hipGraph rounds" if graph else "persistent"},
        "walk_steps_per_s": round(steps_per_s, 1),
        "sites_capped": capped,
        "mean_walks_per_side": round(float(np.mean(res.walks)), 1),
        "rounds": res.stats["rounds"],
        "launch_shape": out_shape,
        "roofline": roofline,
        "results_checksum": checksum,
    }
    if tuning is not None:
        out["autotune"] = {"choice": tuning.choice, "speedup_vs_default_shape": round(tuning.speedup, 4),
                           "trials": len(tuning.trials)}
    if mesh is not None:
        mem = ctx.memory()
        mean_step = total_t / args.steps
        build_s = (mesh.build_ms + grid_build_ms) * 1e-3
        out["lbvh_build_ms"] = round(mesh.build_ms, 3)
        out["geometry"] = {"lbvh_build_ms": round(mesh.build_ms, 3), "grid_build_ms": round(grid_build_ms, 3),
                           "lbvh_bytes": mesh.bvh_bytes, "geometry_bytes": mem["geometry_bytes"],
                           "grid_bytes": mem["geometry_bytes"] - mesh.bvh_bytes,
                           "engine_bytes": mem["engine_bytes"],
                           "sites_per_s_cold": round(float(np.mean(done)) / (mean_step + build_s), 2)}
    if w.precision == 32 and args.fp64_sites > 0:
        out["fp64_rerun"] = fp64_rerun(ctx, geom, sites, sides, kw, shape, args.fp64_sites)
    if not args.no_plain:
        rng = np.random.default_rng(1)
        idx = np.sort(rng.choice(n, size=min(args.plain_sites, n), replace=False))
        t0 = time.perf_counter()
        plain = ctx.estimate(geom, sites[idx], sides, schedule=twos.PLAIN, **kw)
        tp = time.perf_counter() - t0
        plain_rate = finished(plain) / tp
        # The same sample through the batched path, for a like-for-like ratio.
        t0 = time.perf_counter()
        batched = run(twos.PERSISTENT, sites[idx], graph)
        tb = time.perf_counter() - t0
        out["plain_gpu"] = {"sites_per_s": round(plain_rate, 2), "sample_sites": len(idx),
                            "seconds": round(tp, 3), "batched_sample_sites_per_s": round(finished(batched) / tb, 2)}
        out["x_plain"] = round(value / plain_rate, 2) if plain_rate > 0 else None
    if w.container == twos.MESH and not args.no_extras:
        # Configurations 4 and 5 in short form on the same device: 200 requests
        # of the serving stream and two rounds of the design loop.
        t0 = time.perf_counter()
        out["serving"] = serve_run(ctx, 200, 40.0, tuned=not args.no_autotune)
        out["design"] = design_run(ctx, 2, 65536, 4096)
        log(f"serving and design {time.perf_counter() - t0:.1f} s")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, sites, args.cpu_sites)
    ctx.close()
    print(json.dumps(out))
    return 0


def serve_run(ctx, n_requests: int, rate: float, tuned: bool = True) -> dict:
    """Configuration 4 on a context: a Poisson request stream over the three
    resident geometries (20% coupled insertion validations), continuous
    batching, latencies from arrival to the end of the completing round."""
    import twos
    from twos import serving

    stream = serving.Stream(ctx, twos, n_requests=n_requests, rate=rate)
    tuning = None
    if tuned:
        # The mesh requests run at the tuned launch shape and distance grid
        # (the committed choice for this mesh, eps and combine).
        from twos import autotune
        from twos import workloads as W
        w, g = stream.geoms["sh_icosphere"]
        tuning = autotune.autotune(ctx, g, w.sites, W.sides(w, twos.side), geom_sig=w.name, combine=twos.CROSS,
                                   precision=w.precision, eps=w.eps, max_steps=w.max_steps, distance_grid=True,
                                   mesh=stream.mesh, log=log, cache_path=TUNING_CACHE)
        stream.shapes["sh_icosphere"] = {k: v for k, v in tuning.choice.items()
                                         if k not in ("grid_cells", "grid_corners")}
    # Slots and one round's walk buffer; admission charges requests against the
    # context's HBM budget beside the resident geometry.
    srv = ctx.server(max_sites=65536, max_walks_per_round=1 << 26)
    log(f"serving {n_requests} requests at {rate}/s")
    res = serving.run(stream, srv, log=log)
    srv.close()
    out = {"requests": n_requests, "rate_per_s": rate, "p50_ms": round(res["p50_ms"], 2),
           "p99_ms": round(res["p99_ms"], 2), "mean_ms": round(res["mean_ms"], 2),
           "td_sites_per_s": round(res["td_sites_per_s"], 2), "walk_steps_per_s": round(res["walk_steps_per_s"], 1),
           "rounds": res["rounds"], "validations": res["validations"], "td_sites": res["td_sites"],
           "failed": res["failed"], "wall_s": round(res["wall_s"], 3),
           "busy_fraction": round(res["busy_s"] / res["wall_s"], 4)}
    frozen = frozen_unit("sh_icosphere_81920")
    if frozen is not None and res["busy_s"] > 0:
        # The mesh requests' closest-point queries priced at the frozen exact
        # query, over the time spent in server rounds (busy) and over the wall.
        b = res["mesh_queries"] * frozen["frozen_bytes_per_query"]
        out["roofline"] = {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                           "achieved_frozen_busy": round(b / res["busy_s"] / 1e9, 1),
                           "frac_frozen_busy": round(b / res["busy_s"] / 1e9 / HBM_PEAK_GBS, 4),
                           "frac_frozen_wall": round(b / res["wall_s"] / 1e9 / HBM_PEAK_GBS, 4),
                           "mesh_queries": res["mesh_queries"],
                           "basis": "mesh requests' queries x the frozen exact-query bytes (tests/golden/roofline_c3.json)"}
    if tuning is not None:
        out["mesh_launch_shape"] = tuning.choice
    return out


def serve(args) -> int:
    import twos
    ctx = twos.Context()
    res = serve_run(ctx, args.requests, args.rate, tuned=not args.no_autotune)
    ctx.close()
    out = {
        "metric": "serving p50/p99 latency (TD queries + insertion validations)",
        "value": res["p50_ms"], "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(res["wall_s"] * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 walks, f64 combine", "data": "synthetic",
        "config": {"workload": f"serve: {args.requests} Poisson requests at {args.rate}/s over unit ball, cooling "
                               "block and SH icosphere; sites log-uniform 1-8192, CI U[1%,5%], 20% coupled "
                               "insertion validations", "max_sites": 65536, "max_walks_per_round": 1 << 26},
        **res,
    }
    print(json.dumps(out))
    return 0


def design_run(ctx, rounds: int, candidates: int, scout_walks: int) -> dict:
    """Sequential insertion design on the SH-icosphere mesh (configuration 5
    per round): every round scores the feasible candidates (scout, fixed walks
    per side), rescores the best 8 with 16,384 fresh walks per side, accepts
    the best lower confidence bound and inserts a radius-0.02 cavity there;
    then the first and last rounds' rescoring is rerun in binary64 on the same
    Philox streams."""
    import twos
    from twos import autotune
    from twos import design as D
    from twos import workloads as W

    w = W.sh_icosphere(n_sites=candidates)
    mesh = ctx.mesh(w.vertices, w.faces)
    geom = ctx.geometry(twos.MESH, mesh=mesh)
    sides = W.sides(w, twos.side)
    tuning = autotune.autotune(ctx, geom, w.sites, sides, geom_sig=w.name, combine=twos.CROSS, precision=32,
                               eps=w.eps, max_steps=w.max_steps, distance_grid=True, mesh=mesh, log=log,
                               cache_path=TUNING_CACHE)
    shape = {k: v for k, v in tuning.choice.items() if k not in ("grid_cells", "grid_corners")}
    est = D.GpuDesign(ctx, mesh, sides, radius=0.02, eps=w.eps, max_steps=w.max_steps, shape=shape)
    cfg = D.GreedyConfig(rounds=rounds, radius=0.02, clearance=0.025, scout_walks=scout_walks, shortlist=8,
                         rescore_walks=16384, seed=2026, select="lcb")
    # Warm the engine and graphs on one small round, untimed.
    D.greedy(est, w.sites[:2048], D.GreedyConfig(rounds=1, radius=0.02, scout_walks=64, shortlist=2,
                                                 rescore_walks=256, seed=1))
    est.steps = est.paths = 0
    t0 = time.perf_counter()
    out = D.greedy(est, w.sites, cfg, log=log)
    total = time.perf_counter() - t0
    scored = out.candidates_scored
    est64 = D.GpuDesign(ctx, mesh, sides, radius=0.02, eps=w.eps, max_steps=w.max_steps, precision=64, shape=shape)
    rerun = []
    for k in sorted({0, len(out.rounds) - 1}):
        r = out.rounds[k]
        pts = w.sites[r.feasible][r.shortlist]
        s64, se64 = est64(pts, out.centers[:k], cfg.rescore_walks, cfg.seed + cfg.seed_step * k + cfg.rescore_offset)
        s64, se64 = np.asarray(s64), np.asarray(se64)
        rerun.append({"round": k, "sites": len(pts),
                      "max_abs_diff_in_se": float(np.max(np.abs(r.rescore_score - s64) / se64)),
                      "max_rel_diff": float(np.max(np.abs(r.rescore_score - s64) / np.abs(s64)))})
        log(f"[design] binary64 rerun of round {k}: {rerun[-1]}")
    ctx.release_geometry(geom)
    ctx.release_mesh(mesh)
    return {"candidates_per_s": round(scored / total, 2), "rounds": rounds, "candidates": len(w.sites),
            "scout_walks_per_side": scout_walks, "rescore_walks_per_side": 16384, "seconds": round(total, 3),
            "ms_per_round": round(1e3 * total / rounds, 1), "walk_steps_per_s": round(est.steps / total, 1),
            "paths": est.paths,
            "per_round": [{"feasible": len(r.feasible), "chosen": np.round(r.chosen, 6).tolist(),
                           "score": float(r.rescore_score[int(np.argmax(r.rescore_score - 1.96 * r.rescore_se))]),
                           "seconds": round(r.seconds, 3)} for r in out.rounds],
            "binary64_rerun": rerun, "launch_shape": tuning.choice}


def design_bench(args) -> int:
    import twos
    ctx = twos.Context()
    res = design_run(ctx, args.design_rounds, args.design_candidates, args.design_scout_walks)
    ctx.close()
    result = {
        "metric": "design candidates scored per second (scout + rescore)", "value": res["candidates_per_s"],
        "unit": "candidates/s", "n_gpus": 1, "steps": args.design_rounds, "warmup": 1,
        "ms_per_step": res["ms_per_round"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 walks, f64 combine", "data": "synthetic",
        "config": {"workload": f"design on sh_icosphere_81920: {res['candidates']} candidates, {args.design_rounds} "
                               f"rounds, scout {args.design_scout_walks} walks/side, top 8 rescored with 16384 "
                               "fresh walks/side, LCB acceptance, radius-0.02 cavities"},
        **res,
    }
    print(json.dumps(result))
    return 0


COPY_BYTES = 256 << 20


def copy_probe() -> int:
    """A plain device-to-device copy of 256 MiB, three times (the HIP runtime's
    copy kernel): the reading of rocprofv3's FETCH_SIZE / WRITE_SIZE for bytes
    of known count."""
    import ctypes as C
    hip = None
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = C.CDLL(name)
            break
        except OSError:
            continue
    if hip is None:
        raise SystemExit("libamdhip64.so not found")
    a, b = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(COPY_BYTES)
    for call in (lambda: hip.hipMalloc(C.byref(a), n), lambda: hip.hipMalloc(C.byref(b), n),
                 lambda: hip.hipMemset(a, 1, n)):
        if call() != 0:
            raise SystemExit("HIP allocation failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, n, 3) != 0:  # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": COPY_BYTES, "copies": 3}))
    return 0


def shape_args(shape: dict) -> list:
    """bench.py flags that fix a launch shape and distance grid."""
    flags = {"block": "--block", "waves_per_simd": "--waves-per-simd", "lds_stack": "--lds-stack",
             "step_batch": "--step-batch", "refill_batch": "--refill-batch", "grid_kappa": "--grid-kappa",
             "grid_cells": "--grid-cells",
             "grid_corners": "--grid-corners"}
    out = []
    for k, f in flags.items():
        if k in shape:
            out += [f, str(shape[k])]
    return out


def traffic_probe(args) -> int:
    """One eager estimate of the workload at the launch shape given on the
    command line: the walk launches a counter pass reads."""
    import twos
    from twos import workloads as W
    w = build_workload(args.workload)
    sites = w.sites if args.sites <= 0 else w.sites[:args.sites]
    ctx = twos.Context()
    mesh = ctx.mesh(w.vertices, w.faces)
    geom = ctx.geometry(twos.MESH, mesh=mesh)
    grid = args.grid_cells > 0 and not args.exact
    if grid:
        ctx.set_distance_grid(mesh, args.grid_cells, max(args.grid_corners, 0))
    shape = dict(block=args.block, waves_per_simd=max(args.waves_per_simd, 0), lds_stack=args.lds_stack,
                 step_batch=max(args.step_batch, 0), refill_batch=args.refill_batch,
                 grid_kappa=args.grid_kappa if grid else 0.0)
    res = ctx.estimate(geom, sites, W.sides(w, twos.side), combine=twos.CROSS, precision=w.precision, seed=2026,
                       walks=w.walks, walks_max=w.walks_max, ci_rel=w.ci_rel, eps=w.eps, max_steps=w.max_steps,
                       distance_grid=grid, **shape)
    ctx.close()
    print(json.dumps({"traffic_probe": True, "walk_launches": res.stats["walk_launches"]}), flush=True)
    return 0


def _pmc_pass(counter: str, program: list, timeout: float) -> list:
    """(kernel name, counter value) per dispatch of one rocprofv3 --pmc pass
    over `program` (a child process; the counter CSV is read even when the
    profiler itself exits non-zero after the program finished)."""
    import csv
    import subprocess
    import tempfile
    out = Path(tempfile.mkdtemp(prefix="twos_pmc_", dir="/tmp"))
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", str(out), "-o", "run", "--", *program]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd="/tmp", env=env)
    files = sorted(out.rglob("*counter_collection.csv"))
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter file (rc {r.returncode}): {r.stderr[-300:]}")
    rows = []
    for f in files:
        with f.open() as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") == counter:
                    rows.append((row["Kernel_Name"], float(row["Counter_Value"])))
    return rows


def live_traffic(workload: str, n_sites: int, shape: dict, launches: int):
    """HBM bytes per walk-kernel launch measured in this run: two rocprofv3
    --pmc passes (FETCH_SIZE, WRITE_SIZE, KiB per dispatch) over the same
    estimate at the same shape, each scaled by what the counter reads for the
    256-MiB device-to-device copy (two more passes).  Returns (bytes, source)
    or (None, why not)."""
    import shutil
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not on PATH"
    probe = [sys.executable, str(ROOT / "bench.py"), "--workload", "sh_icosphere", "--traffic-probe",
             "--sites", str(n_sites), *shape_args(shape), *([] if "grid_cells" in shape else ["--exact"])]
    copy = [sys.executable, str(ROOT / "bench.py"), "--workload", "copy"]
    try:
        per = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            t0 = time.perf_counter()
            walk = [v for k, v in _pmc_pass(counter, probe, 240) if "walk_persistent<float" in k]
            cal = [v for k, v in _pmc_pass(counter, copy, 120) if "copyBuffer" in k]
            log(f"traffic pass {counter}: {len(walk)} walk launches, {len(cal)} copies, "
                f"{time.perf_counter() - t0:.1f} s")
            if len(walk) != launches or not cal:
                return None, f"{counter}: {len(walk)} walk launches (expected {launches}), {len(cal)} copies"
            scale = (COPY_BYTES / 1024) / float(np.median(cal))  # KiB of data per KiB counted
            per[counter] = sum(walk) * scale * 1024 / launches
        return int(per["FETCH_SIZE"] + per["WRITE_SIZE"]), "rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE in this run"
    except Exception as e:  # noqa: BLE001 - a profiler failure must not fail the benchmark
        return None, f"live measurement failed: {e}"[:300]


def frozen_unit(workload: str):
    path = ROOT / "tests" / "golden" / "roofline_c3.json"
    if not path.exists():
        return None
    d = json.loads(path.read_text())
    return d if d.get("workload") == workload else None


def fp64_rerun(ctx, geom, sites, sides, kw, shape, count: int) -> dict:
    """The binary32 walks' error stated against a binary64 GPU rerun of the
    same Philox streams (same shape, grid and threshold): every count-th
    site at a fixed 4096 walks per side."""
    import twos
    sub = sites[:: max(1, len(sites) // count)][:count]
    fixed = {**kw, "walks": 4096, "walks_max": 0}
    r32 = ctx.estimate(geom, sub, sides, schedule=twos.PERSISTENT, **{**fixed, "precision": 32}, **shape)
    r64 = ctx.estimate(geom, sub, sides, schedule=twos.PERSISTENT, **{**fixed, "precision": 64}, **shape)
    dev = np.abs(r32.score - r64.score) / r64.score_se
    return {"sites": len(sub), "walks_per_side": 4096, "max_abs_diff_in_se": round(float(np.max(dev)), 4),
            "median_abs_diff_in_se": round(float(np.median(dev)), 5),
            "max_rel_diff": float(np.max(np.abs(r32.score - r64.score) / np.abs(r64.score))),
            "steps_max_rel_diff": float(np.max(np.abs(r32.steps - r64.steps) / r64.steps))}


def measured_traffic(workload: str, n_sites: int, shape: dict, launches: int):
    """HBM bytes per walk-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by a copy kernel of known
    size).  The estimate is deterministic, so the same workload and launch
    shape repeat the same launches; any other configuration reports null."""
    path = ROOT / "profiles" / "walk_persistent_sh_icosphere_16384.md"
    if workload != "sh_icosphere_81920" or n_sites != 16384 or not path.exists():
        return None
    fields = {}
    for line in path.read_text().splitlines():
        if ":" in line and line.split(":")[0] in ("traffic_bytes_per_launch", "traffic_launches", "traffic_shape"):
            key, val = line.split(":", 1)
            fields[key] = val.strip()
    if "traffic_shape" not in fields or json.loads(fields["traffic_shape"]) != shape:
        return None
    if float(fields.get("traffic_launches", -1)) != launches:
        return None
    return int(float(fields["traffic_bytes_per_launch"]))


def _diff_z(a, b):
    """z-scores of two Differential WoS results' derivative columns in their
    combined SE (entries with a positive SE)."""
    zs = []
    for col, se in ((0, 3), (1, 4), (2, 5), (6, 7)):
        s_ = np.hypot(a[..., se], b[..., se])
        ok = s_ > 0
        zs.append((a[..., col] - b[..., col])[ok] / s_[ok])
    return np.concatenate(zs)


def f4_bench(args) -> int:
    """The adjacent estimators on the SH-icosphere mesh (81,920 triangles):
    Differential WoS (twos_differential; 1,024 sites x 1,024 samples, four
    cavities of radius 0.05, the primal data g = x + 0.3 y^2, offset 0.004,
    eps 1e-4) or certified minimum-benefit sampling (twos_certify; 16 sites x
    65,536 samples, the affine source, radius cap 0.1, value cap 0.06, eps
    1e-6, 128 steps).  A step is one call; the metric is samples per second.
    Certified sampling is timed in its production form (Philox streams, one
    persistent launch, binary64 with directed rounding, the mesh's 256-cell
    certified grid for steps at least two cells from the mesh), its replay
    form (the reference's mt19937_64 streams) and the exact Philox form once
    beside it.  Differential WoS is timed
    in its production form (Philox streams, binary32 walks, the mesh's 1024-cell
    record grid for steps at least 2 eps clear of the mesh); its replay form
    (binary64, the reference's mt19937_64 streams) and the other production
    variants are timed once each beside it, with the production result's
    agreement with the exact binary64 Philox walk in combined SEs."""
    import twos
    from twos import workloads as W

    w = W.sh_icosphere(n_sites=1024 + 16)
    ctx = twos.Context()
    mesh = ctx.mesh(w.vertices, w.faces)
    c = w.sites[:4]
    diff = args.workload == "diffwos"
    geom = ctx.geometry(twos.MESH, mesh=mesh, cavities=c, cavity_radius=0.05)
    forms = {}
    if diff:
        sites = w.sites[16:16 + 1024]
        sites = sites[np.min(np.linalg.norm(sites[:, None] - c[None], axis=2), axis=1) > 0.06]
        side = W.sides(w, twos.side)[0]
        walks = 1024

        def run(**kw):
            return ctx.differential(geom, sites, side, seed=11, walks=walks, eps=1e-4, max_steps=2048,
                                    offset=0.004, **kw)

        def once(name, **kw):
            run(**kw)
            t0 = time.perf_counter()
            out = run(**kw)
            dt = time.perf_counter() - t0
            st = int(out[1][:, 0].sum() + out[1][:, 1].sum())
            forms[name] = {"samples_per_s": round(len(sites) * walks / dt, 1), "walk_steps_per_s": round(st / dt, 1),
                           "ms": round(1e3 * dt, 2)}
            log(f"{name}: {dt:.3f} s, {st} steps")
            return out

        once("replay_f64_mt19937")
        exact = once("philox_f64", rng=twos.RNG_PHILOX, precision=64)
        once("philox_f32", rng=twos.RNG_PHILOX, precision=32)
        once("philox_f32_grid128_centres", rng=twos.RNG_PHILOX, precision=32, distance_grid=True)
        grid_ms = ctx.set_distance_grid(mesh, 1024, 2)
        log(f"1024-cell record grid built in {grid_ms:.0f} ms")
        prod = dict(rng=twos.RNG_PHILOX, precision=32, distance_grid=True, grid_kappa=0.0625)
        call = lambda: run(**prod)
    else:
        sites = w.sites[4:16]
        sites = sites[np.min(np.linalg.norm(sites[:, None] - c[None], axis=2), axis=1) > 0.06]
        walks = 65536

        def cert(philox, grid_cells=0):
            return ctx.certify(geom, sites, seed=11, walks=walks, max_steps=128, eps=1e-6, radius_cap=0.1,
                               value_cap=0.06, source=twos.CERT_AFFINE, slope=0.65, philox=philox,
                               grid_cells=grid_cells)

        def cert_once(name, philox, grid_cells=0):
            cert(philox, grid_cells)
            t0 = time.perf_counter()
            out_ = cert(philox, grid_cells)
            dt = time.perf_counter() - t0
            forms[name] = {"samples_per_s": round(len(sites) * walks / dt, 1),
                           "walk_steps_per_s": round(float(out_[1].sum()) / dt, 1), "ms": round(1e3 * dt, 2)}
            log(f"{name}: {dt:.3f} s")
            return out_

        replay = cert_once("replay_f64_mt19937", False)
        cert_once("philox_f64_exact", True)
        cert_once("philox_f64_grid128", True, 128)
        cert_once("philox_f64_grid512", True, 512)
        t0 = time.perf_counter()
        ctx.cert_grid_lower(mesh, 256, sites[:1])
        forms["cert_grid_build_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
        call = lambda: cert(True, 256)
    for i in range(args.warmup):
        call()
        log(f"warmup {i}")
    times, steps = [], 0
    for i in range(args.steps):
        t0 = time.perf_counter()
        out = call()
        times.append(time.perf_counter() - t0)
        steps = int(out[1][:, 0].sum() + out[1][:, 1].sum()) if diff else int(out[1].sum())
        log(f"step {i}: {times[-1]:.3f} s, {steps} steps")
    t = float(np.mean(times))
    samples = len(sites) * walks
    result = {
        "metric": ("Differential WoS samples/s" if diff else "certified samples/s"), "value": round(samples / t, 1),
        "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * t, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"diffwos on {w.name}: {len(sites)} sites x {walks} samples, 4 cavities r 0.05"
                                if diff else f"certify on {w.name}: {len(sites)} sites x {walks} samples")},
        "walk_steps_per_s": round(steps / t, 1),
        # Identical across launch shapes and batching choices (A/B check).
        "results_checksum": hashlib.sha256(b"".join(np.ascontiguousarray(a).tobytes() for a in out)).hexdigest()[:16],
    }
    if diff:
        z = _diff_z(out[0], exact[0])
        result["dtype"] = "f32 walks, f64 fold"
        result["config"]["form"] = "Philox streams, binary32, 1024-cell record grid, threshold 2 eps"
        result["forms"] = forms
        result["x_replay"] = round(result["value"] / forms["replay_f64_mt19937"]["samples_per_s"], 1)
        result["vs_exact_philox_f64"] = {"entries": int(len(z)), "max_abs_z": round(float(np.max(np.abs(z))), 3),
                                         "mean_z": round(float(np.mean(z)), 4),
                                         "beyond_3_se": int(np.sum(np.abs(z) > 3))}
    else:
        # Per site: the Philox and the reference-stream means of the certified
        # value in their combined SE, and both forms' containment failures.
        va, vb = out[0], replay[0]
        se = np.hypot(va.std(axis=1), vb.std(axis=1)) / np.sqrt(walks)
        z = (va.mean(axis=1) - vb.mean(axis=1)) / np.where(se > 0, se, np.inf)
        result["config"]["form"] = ("Philox streams, binary64 with directed rounding, one persistent launch, "
                                    "256-cell certified grid for steps at least 2 cells from the mesh")
        result["forms"] = forms
        result["x_replay"] = round(result["value"] / forms["replay_f64_mt19937"]["samples_per_s"], 1)
        result["vs_replay"] = {"sites": int(len(z)), "max_abs_z": round(float(np.max(np.abs(z))), 3),
                               "containment_failures": int(np.sum(out[2] & twos.CERT_CONTAINMENT)
                                                           + np.sum(replay[2] & twos.CERT_CONTAINMENT))}
    ctx.release_geometry(geom)
    ctx.close()
    print(json.dumps(result))
    return 0


def cpu_baseline(w, sites, count: int, repeats: int = 3) -> dict:
    """The CPU oracle (binary64, same Philox streams, same ladder) on a seeded
    sample of the sites, OpenMP over every host core the process may use;
    the median of `repeats` timed runs, as the reference reports its timings
    (wih_twos.cpp:1751)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from twos import workloads as W

    cores = O.host_threads()
    rng = np.random.default_rng(2)
    idx = np.sort(rng.choice(len(sites), size=min(count, len(sites)), replace=False))
    if w.container == 2:
        geo = O.Geometry(O.MESH, mesh=O.Mesh(w.vertices, w.faces))
    elif w.container == 1:
        geo = O.Geometry(O.BOX, box_lo=w.box_lo, box_hi=w.box_hi, spheres=w.spheres, cylinders=w.cylinders)
    else:
        geo = O.Geometry(O.BALL, ball=w.ball)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        r = O.estimate(geo, sites[idx], W.sides(w, O.side), combine=O.CROSS, rng=O.RNG_PHILOX, seed=2026,
                       walks=w.walks, walks_max=w.walks_max, ci_rel=w.ci_rel, eps=w.eps,
                       max_steps=w.max_steps, threads=cores)
        times.append(time.perf_counter() - t0)
        log(f"cpu baseline run {len(times)}: {times[-1]:.1f} s")
    t = float(np.median(times))
    if w.walks_max > w.walks:
        ok = int(np.sum((r["status"] == 0) & (1.96 * r["score_se"] <= w.ci_rel * np.abs(r["score"]))))
    else:
        ok = len(idx)
    return {"value": round(ok / t, 3), "unit": "sites/s", "cores": cores, "kind": "port",
            "sample": f"{len(idx)} seeded sites of the same workload, median of {repeats} runs {t:.1f} s, "
                      f"{int(r['steps'].sum())} steps ({r['steps'].sum() / t:.3e} steps/s)"}


if __name__ == "__main__":
    sys.exit(main())
