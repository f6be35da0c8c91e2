"""Autotuner for the persistent walk kernel's launch choices.

Axes (binary32 persistent kernel; binary64 tunes the block and the refill
batch only):
  block           64, 128 or 256 lanes per block
  waves_per_simd  0 = the compiler's register allocation (4 waves per SIMD),
                  5, 6 or 8 (registers capped for that occupancy; 6 and 8
                  spill to scratch)
  lds_stack       16, 24 or 32 LDS traversal-stack entries per lane (16 and
                  32 exist for waves_per_simd 0 and 5)
  step_batch      0 (a lane takes its walk step as soon as its query ends),
                  1 (once at least as many lanes are ready as traverse; the
                  node visits wait until the traversing lanes outnumber the
                  ready ones) or 32
                  (once 32 lanes are ready); built for lds_stack 24 with
                  waves_per_simd 0 or 5
  refill_batch    1, 4, 8, 16 or 32: a wave's idle lanes take new paths once
                  that many are idle (binary32 and binary64)
  grid_kappa      distance-grid threshold in cell sizes, 0.0625 to 6 (mesh
                  geometries with the distance grid on)
  grid_corners    distance-grid samples at cell centres (0), corners (1,
                  eight-corner bound) or binary16 corner records per cell (2),
                  with the mesh handle
  grid_cells      distance-grid resolution, 96 to 1024 cells along
                  the mesh's longest axis (when the mesh handle is given; the
                  mesh keeps the chosen grid)

The launch shape never changes a result: every path is keyed by (seed, site,
side, walk) and each site's samples are folded in walk order.  grid_kappa
changes paths (a different mix of grid and exact radii) but not the
estimator's expectation, so every axis is scored by the same objective: walk
kernel time (HIP events) per path, on a fixed-budget sample of the caller's
sites.

Search: coordinate descent from the defaults (256, 0, 24, 4), one axis at a
time, twice over, each candidate the best of `repeats` timed runs; a
candidate replaces the current choice only when it is at least 1% faster
(run-to-run noise of the kernel time is below that).  Results
are cached in memory per (device, geometry signature, precision, combine,
eps, distance grid) and, when `cache_path` is given, in a JSON file.
"""

from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _capi

DEFAULTS = {"block": 256, "waves_per_simd": 0, "lds_stack": 24, "step_batch": 0, "refill_batch": 8, "grid_corners": 0,
            "grid_kappa": 4.0, "grid_cells": 128}
AXES = {
    "block": (64, 128, 256),
    "waves_per_simd": (0, 5, 6, 8),
    "lds_stack": (16, 24, 32),
    "step_batch": (0, 1, 32),
    "refill_batch": (1, 4, 8, 16, 32),
    "grid_corners": (0, 1, 2),
    "grid_kappa": (0.0625, 0.125, 0.25, 0.35, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0),
    "grid_cells": (96, 128, 192, 256, 320, 384, 448, 512, 640, 768, 896, 1024),
}

_memory: dict[str, "Tuning"] = {}


@dataclass
class Tuning:
    choice: dict
    ms_per_path: float
    default_ms_per_path: float
    trials: list = field(default_factory=list)  # (config, ms per path)
    grid_build_ms: float = 0.0  # build of the chosen distance grid when it was last applied to the mesh

    @property
    def speedup(self) -> float:
        return self.default_ms_per_path / self.ms_per_path

    def as_dict(self) -> dict:
        return {"choice": self.choice, "ms_per_path": self.ms_per_path,
                "default_ms_per_path": self.default_ms_per_path, "speedup": self.speedup,
                "trials": [[c, t] for c, t in self.trials]}


def _valid(cfg: dict, precision: int) -> bool:
    batch = cfg.get("step_batch", 0)
    if precision == 64:
        return cfg["waves_per_simd"] == 0 and cfg["lds_stack"] == 24 and batch == 0
    if cfg["lds_stack"] != 24 and cfg["waves_per_simd"] not in (0, 5):
        return False
    if batch and (cfg["lds_stack"] != 24 or cfg["waves_per_simd"] not in (0, 5)):
        return False
    return True


def _key(ctx, geom_sig: str, precision: int, combine: int, eps: float, grid: bool) -> str:
    return f"dev{ctx.device}|{geom_sig}|p{precision}|c{combine}|e{eps:g}|g{int(grid)}"


def autotune(ctx, geom: int, sites, sides, *, geom_sig: str, combine: int, precision: int = 32,
             eps: float = 1e-4, max_steps: int = 8192, distance_grid: bool = False, mesh=None,
             sample: int = 4096, walks: int = 512, seed: int = 12345, repeats: int = 2,
             cache_path: str | os.PathLike | None = None, log=None) -> Tuning:
    """Choose the launch shape for estimates like the one described.

    geom_sig names the geometry for the cache (for example the mesh's triangle
    count and kind); sites/sides are representative of the calls to tune for
    (a sample of min(len(sites), sample) sites, evenly spaced, runs `walks`
    walks per side with no adaptivity).  With `mesh` (the geometry's mesh
    handle) and the distance grid on, the grid resolution is tuned too and
    left at the chosen value."""
    key = _key(ctx, geom_sig, precision, combine, eps, distance_grid) + ("|m" if mesh is not None else "")
    tune_cells = mesh is not None and distance_grid

    def applied(t: "Tuning") -> "Tuning":
        if tune_cells:
            t.grid_build_ms = ctx.set_distance_grid(mesh, t.choice["grid_cells"], int(t.choice["grid_corners"]))
        return t

    if key in _memory:
        return applied(_memory[key])
    if cache_path is not None and Path(cache_path).exists():
        stored = json.loads(Path(cache_path).read_text())
        if key in stored:
            d = stored[key]
            t = Tuning(d["choice"], d["ms_per_path"], d["default_ms_per_path"], [tuple(x) for x in d["trials"]])
            _memory[key] = t
            return applied(t)
    pts = np.asarray(sites, dtype=np.float64).reshape(-1, 3)
    stride = max(1, len(pts) // sample)
    pts = pts[::stride][:sample]
    axes = dict(AXES)
    if not distance_grid:
        axes.pop("grid_kappa")
    if not tune_cells:
        axes.pop("grid_cells")
        axes.pop("grid_corners")
    if precision == 64:
        axes = {k: v for k, v in axes.items()
                if k in ("block", "refill_batch", "grid_kappa", "grid_cells", "grid_corners")}
    measured: dict[tuple, float] = {}
    grid_now = [None]

    def cost(cfg: dict) -> float:
        k = tuple(sorted(cfg.items()))
        if k in measured:
            return measured[k]
        if tune_cells and (cfg["grid_cells"], cfg["grid_corners"]) != grid_now[0]:
            ctx.set_distance_grid(mesh, cfg["grid_cells"], int(cfg["grid_corners"]))
            grid_now[0] = (cfg["grid_cells"], cfg["grid_corners"])
        best = math.inf
        for _ in range(repeats):
            r = ctx.estimate(geom, pts, sides, combine=combine, precision=precision, seed=seed, walks=walks,
                             eps=eps, max_steps=max_steps, distance_grid=distance_grid,
                             block=cfg["block"], waves_per_simd=cfg["waves_per_simd"], lds_stack=cfg["lds_stack"],
                             step_batch=cfg.get("step_batch", 0), refill_batch=cfg.get("refill_batch", 0),
                             grid_kappa=cfg["grid_kappa"] if distance_grid else 0.0, raise_on_cap=False)
            best = min(best, r.stats["walk_kernel_ms"] / max(1, r.stats["paths"]))
        measured[k] = best
        if log:
            log(f"[autotune] {cfg} -> {best * 1e6:.2f} ns/path")
        return best

    cfg = {k: v for k, v in DEFAULTS.items() if k in axes or k in ("block", "waves_per_simd", "lds_stack",
                                                                   "grid_kappa")}
    default_cost = cost(cfg)
    for _ in range(2):
        changed = False
        for axis, values in axes.items():
            for v in values:
                trial = {**cfg, axis: v}
                if trial == cfg or not _valid(trial, precision):
                    continue
                if cost(trial) < cost(cfg) * 0.99:
                    cfg = trial
                    changed = True
        if not changed:
            break
    t = Tuning(cfg, cost(cfg), default_cost, [(dict(k), v) for k, v in measured.items()])
    applied(t)
    _memory[key] = t
    if cache_path is not None:
        path = Path(cache_path)
        stored = json.loads(path.read_text()) if path.exists() else {}
        stored[key] = t.as_dict()
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(stored, indent=1) + "\n")
    return t


def clear_cache() -> None:
    _memory.clear()
