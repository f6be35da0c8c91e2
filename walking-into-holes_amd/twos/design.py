"""Sequential insertion design and insertion validation: the estimator's main
callers (benchmarks/submission_study/run_study.py:77-106,
benchmarks/internal_cooling/run_reference_benchmark.py:183-245,
benchmarks/stanford_bunny/run_bunny_benchmark.py:290-337).

Greedy insertion, one cavity per round:
  1. feasible candidates: farther than 2 r + clearance from every cavity
     already placed (run_study.py:85-87);
  2. scout: estimate every feasible candidate's benefit with few walks;
  3. shortlist: the best `shortlist` scout scores (a stable sort, so ties keep
     candidate order; the reference's numpy argsort is not stable, and which
     of several equal scores it keeps depends on the numpy build);
  4. rescore the shortlist with fresh walks (a new seed) and accept the best
     mean, or the best lower confidence bound score - z se;
  5. the accepted centre becomes a cavity of radius r for the next round.

Insertion validation (`validate_insertion`): the objective J = integral q u,
q the unit-mass sensor bump, estimated at sensor-distributed probes, on the
domain with and without the candidate ball; the two estimates use the same
streams (the same seed and probe index), so their paths coincide until the
ball changes them, and their difference over the ball's capacity 4 pi r
(benchmarks/submission_study/README.md:27-34) is the small-ball benefit the
topological derivative predicts.

The estimator is a callable, so the same loop drives the GPU (`GpuStudy`) and
the CPU oracle in tests.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class GreedyConfig:
    rounds: int = 6
    radius: float = 0.065           # cavity radius (run_study.py:10)
    clearance: float = 0.025        # feasibility gap beyond 2 r (run_study.py:86)
    scout_walks: int = 48
    shortlist: int = 16
    rescore_walks: int = 1024
    seed: int = 1611
    seed_step: int = 101            # round k scouts with seed + 101 k (run_study.py:92)
    rescore_offset: int = 53        # and rescores with seed + 101 k + 53 (:94)
    select: str = "mean"            # "mean" (the reference) or "lcb"
    lcb_z: float = 1.96


@dataclass
class Round:
    index: int
    feasible: np.ndarray            # indices into the pool
    scout_score: np.ndarray
    shortlist: np.ndarray           # indices into `feasible`
    rescore_score: np.ndarray
    rescore_se: np.ndarray
    chosen: np.ndarray              # centre accepted this round
    seconds: float


@dataclass
class Design:
    centers: np.ndarray
    rounds: list = field(default_factory=list)

    @property
    def seconds(self) -> float:
        return float(sum(r.seconds for r in self.rounds))

    @property
    def candidates_scored(self) -> int:
        return int(sum(len(r.feasible) + len(r.shortlist) for r in self.rounds))


def feasible(pool: np.ndarray, holes: np.ndarray, radius: float, clearance: float) -> np.ndarray:
    """Indices of pool points farther than 2 r + clearance from every hole."""
    pool = np.asarray(pool, dtype=np.float64).reshape(-1, 3)
    holes = np.asarray(holes, dtype=np.float64).reshape(-1, 3)
    if len(holes) == 0:
        return np.arange(len(pool))
    d = np.min(np.linalg.norm(pool[:, None, :] - holes[None, :, :], axis=2), axis=1)
    return np.nonzero(d > 2 * radius + clearance)[0]


def greedy(estimate, pool, cfg: GreedyConfig, holes=(), log=None, shortlist_fn=None) -> Design:
    """Greedy insertion.  estimate(points, holes, walks, seed) returns
    (score, se) per point: the benefit estimate of a cavity there given the
    cavities `holes` already placed.  shortlist_fn(round, scout_scores), when
    given, returns the shortlist indices (for example to follow a recorded
    run's order among equal scores)."""
    pool = np.asarray(pool, dtype=np.float64).reshape(-1, 3)
    placed = [np.asarray(h, dtype=np.float64) for h in np.asarray(holes, dtype=np.float64).reshape(-1, 3)]
    rounds = []
    for k in range(cfg.rounds):
        t0 = time.perf_counter()
        a = np.array(placed).reshape(-1, 3)
        idx = feasible(pool, a, cfg.radius, cfg.clearance)
        if len(idx) == 0:
            raise ValueError(f"round {k}: no feasible candidate left")
        p = pool[idx]
        seed = cfg.seed + cfg.seed_step * k
        scout, _ = estimate(p, a, cfg.scout_walks, seed)
        if shortlist_fn is not None:
            short = np.asarray(shortlist_fn(k, np.asarray(scout)))
        else:
            short = np.argsort(np.asarray(scout), kind="stable")[-min(cfg.shortlist, len(p)):]
        rescore, se = estimate(p[short], a, cfg.rescore_walks, seed + cfg.rescore_offset)
        rescore, se = np.asarray(rescore), np.asarray(se)
        key = rescore if cfg.select == "mean" else rescore - cfg.lcb_z * np.nan_to_num(se, nan=np.inf)
        best = int(short[int(np.argmax(key))])
        placed.append(p[best])
        rounds.append(Round(k, idx, np.asarray(scout), short, rescore, se, p[best].copy(),
                            time.perf_counter() - t0))
        if log:
            log(f"[design] round {k}: {len(p)} feasible, chose {np.round(p[best], 4).tolist()} "
                f"score {rescore[int(np.argmax(key))]:.4e}")
    return Design(np.array(placed[len(np.asarray(holes).reshape(-1, 3)):]), rounds)


def ball_probes(center, radius: float, n: int, seed: int):
    """Probe points uniform in the sensor ball and their weights
    (1 - rho^2)^2 35/8, so mean(w u) estimates integral q u for the unit-mass
    bump q (the weights of benchmarks/submission_study/geometry.py:82-87; the
    points here come from a seeded numpy generator)."""
    g = np.random.default_rng(seed)
    u = g.random((n, 3))
    z = 1 - 2 * u[:, 0]
    phi = 2 * np.pi * u[:, 1]
    rho = u[:, 2] ** (1 / 3)
    s = np.sqrt(np.maximum(0.0, 1 - z * z))
    p = np.stack([s * np.cos(phi), s * np.sin(phi), z], 1) * rho[:, None]
    return p * radius + np.asarray(center, dtype=np.float64), (1 - rho * rho) ** 2 * 35 / 8


@dataclass
class Validation:
    delta_j: float          # J(domain) - J(domain minus the ball)
    se: float
    per_capacity: float     # delta_j / (4 pi r), comparable to the TD score u p
    per_capacity_se: float
    j_base: float
    j_with: float


def validate_insertion(estimate_u, probes, weights, base_holes, center, radius: float, walks: int,
                       seed: int) -> Validation:
    """Fresh-seed validation of one insertion.  estimate_u(points, holes,
    walks, seed) returns the primal mean u at each point; both calls use the
    same seed, so each probe's paths are coupled."""
    holes = np.asarray(base_holes, dtype=np.float64).reshape(-1, 3)
    w = np.asarray(weights, dtype=np.float64)
    u0 = np.asarray(estimate_u(probes, holes, walks, seed))
    u1 = np.asarray(estimate_u(probes, np.vstack([holes, np.asarray(center, dtype=np.float64)]), walks, seed))
    d = w * (u0 - u1)
    n = len(d)
    delta = float(np.mean(d))
    se = float(np.std(d, ddof=1) / math.sqrt(n)) if n > 1 else float("nan")
    cap = 4 * math.pi * radius
    return Validation(delta, se, delta / cap, se / cap, float(np.mean(w * u0)), float(np.mean(w * u1)))


class GpuDesign:
    """Estimator adapter for any primal/adjoint pair of sides on a resident
    mesh: cavities of one radius, CROSS on Philox streams (binary32 walks by
    default), an optional launch shape from the autotuner.  Returns the
    cross-index score u p and its jackknife SE per point."""

    def __init__(self, ctx, mesh, sides, *, radius: float, eps: float, max_steps: int, precision: int = 32,
                 distance_grid: bool = True, shape: dict | None = None):
        import twos
        self.twos, self.ctx, self.mesh, self.sides = twos, ctx, mesh, sides
        self.radius, self.eps, self.max_steps, self.precision = radius, eps, max_steps, precision
        self.distance_grid, self.shape = distance_grid, dict(shape or {})
        self.steps = 0
        self.paths = 0

    def __call__(self, points, holes, walks, seed):
        t = self.twos
        g = self.ctx.geometry(t.MESH, mesh=self.mesh, cavities=holes if len(holes) else None,
                              cavity_radius=self.radius)
        try:
            r = self.ctx.estimate(g, points, self.sides, combine=t.CROSS, precision=self.precision, seed=seed,
                                  walks=walks, eps=self.eps, max_steps=self.max_steps,
                                  distance_grid=self.distance_grid, use_graph=True, **self.shape)
        finally:
            self.ctx.release_geometry(g)
        self.steps += int(r.stats["steps"])
        self.paths += int(r.stats["paths"])
        return r.score, r.score_se


class GpuStudy:
    """Estimator adapter for the study problem on the GPU: heat bumps and the
    unit-mass sensor bump at a Green point, cavities of one radius.

    streams="reference" replays the reference's paired streams in binary64
    (PAIRED, mt19937_64 keyed as paired_site keys them, study.cpp:89);
    streams="philox" runs the production path (CROSS primal x adjoint on
    independent Philox streams, binary32 walks unless precision=64)."""

    def __init__(self, ctx, mesh, config_text: str, *, radius: float, eps: float = 2e-4, max_steps: int = 2048,
                 streams: str = "reference", precision: int = 32, distance_grid: bool = False):
        import twos
        self.twos, self.ctx, self.mesh = twos, ctx, mesh
        heat, sensor = twos.study_terms(config_text)
        self.heat, self.sensor = twos.side(terms=heat), twos.side(terms=sensor)
        self.radius, self.eps, self.max_steps = radius, eps, max_steps
        self.streams, self.precision, self.distance_grid = streams, precision, distance_grid
        self.steps = 0

    def _run(self, points, holes, walks, seed, hot):
        t = self.twos
        g = self.ctx.geometry(t.MESH, mesh=self.mesh, cavities=holes if len(holes) else None,
                              cavity_radius=self.radius)
        try:
            sides = [self.heat, t.side() if hot else self.sensor]
            if self.streams == "reference":
                r = self.ctx.estimate(g, points, sides, combine=t.PAIRED, rng=t.RNG_MT_PAIRED, precision=64,
                                      seed=seed, walks=walks, eps=self.eps, max_steps=self.max_steps)
            else:
                r = self.ctx.estimate(g, points, sides, combine=t.PAIRED if hot else t.CROSS,
                                      precision=self.precision, seed=seed, walks=walks, eps=self.eps,
                                      max_steps=self.max_steps, distance_grid=self.distance_grid,
                                      use_graph=True)
        finally:
            self.ctx.release_geometry(g)
        self.steps += int(r.stats["steps"])
        return r

    def __call__(self, points, holes, walks, seed):
        r = self._run(points, holes, walks, seed, hot=False)
        # The benefit per capacity is u p: the paired or cross-index score.
        return r.score, r.score_se

    def u(self, points, holes, walks, seed):
        return self._run(points, holes, walks, seed, hot=True).u
