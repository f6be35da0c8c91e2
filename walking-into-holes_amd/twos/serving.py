"""Request stream over the continuous-batching server (configuration 4).

Resident geometries: the unit ball, the cooling block and the SH icosphere.
Requests arrive by a Poisson process; each has a log-uniform number of sites
in [1, max_sites], a CI target U[ci_lo, ci_hi] of |TD| and its own seed.  A
fraction are fresh-seed insertion validations: a sphere of radius r ~ U[0.01,
0.05] is cut out of the geometry at a previously scored site z, and the
change of the objective J = integral q u, q the unit-mass sensor bump on the
ball B(z, 3 r), is re-estimated by the coupled finite-cavity difference
(COUPLED: each walk starts at a q-distributed point and runs, on one stream,
once in the geometry and once in the geometry minus the sphere), as
J(domain) - J(domain minus the sphere) over the sphere's capacity 4 pi r
(benchmarks/submission_study/study.cpp:317-332; capacity,
benchmarks/submission_study/README.md:27-34).

Requests are admitted between rounds whenever the server has free slots, and
every admitted site of every request shares each persistent launch.  Latency
is measured from the arrival time to the end of the round that completes the
request, on the host clock.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import workloads as W

VALIDATION_WALKS = 4096
SENSOR_RADII = 3.0  # the validation's sensor ball, in cavity radii


@dataclass
class Request:
    rid: int
    arrival: float
    kind: str                 # "td" or "validation"
    geom_name: str
    sites: np.ndarray
    seed: int
    ci_rel: float
    geom: int = -1
    ids: list = field(default_factory=list)     # server request ids
    done: int = 0
    results: list = field(default_factory=list)
    finish: float = math.nan
    extra: dict = field(default_factory=dict)


class Stream:
    """Builds the resident geometries and a seeded request list."""

    def __init__(self, ctx, twos_mod, n_requests: int, rate: float, seed: int = 11, max_sites: int = 8192,
                 validation_frac: float = 0.2, ci_lo: float = 0.01, ci_hi: float = 0.05,
                 walks: int = 256, walks_max: int = 16384):
        self.ctx, self.twos = ctx, twos_mod
        self.walks, self.walks_max = walks, walks_max
        # Launch shape and grid threshold per geometry name (for example the
        # autotuner's choice); neither changes which requests finish when.
        self.shapes: dict[str, dict] = {}
        ball = W.unit_ball()
        block = W.cooling_block()
        ico = W.sh_icosphere()
        mesh = ctx.mesh(ico.vertices, ico.faces)
        self.mesh = mesh
        self.geoms = {
            "unit_ball": (ball, ctx.geometry(twos_mod.BALL)),
            "cooling_block": (block, ctx.geometry(twos_mod.BOX, box_lo=block.box_lo, box_hi=block.box_hi,
                                                  spheres=block.spheres, cylinders=block.cylinders)),
            "sh_icosphere": (ico, ctx.geometry(twos_mod.MESH, mesh=mesh)),
        }
        rng = np.random.default_rng(seed)
        names = list(self.geoms)
        t = 0.0
        self.requests: list[Request] = []
        scored: dict[str, list] = {k: [] for k in names}
        for i in range(n_requests):
            t += rng.exponential(1.0 / rate)
            name = names[int(rng.integers(len(names)))]
            w = self.geoms[name][0]
            count = int(round(math.exp(rng.uniform(0.0, math.log(max_sites)))))
            count = max(1, min(count, max_sites))
            is_val = rng.uniform() < validation_frac and len(scored[name]) > 0
            if is_val:
                z = scored[name][int(rng.integers(len(scored[name])))]
                r = float(rng.uniform(0.01, 0.05))
                # Every start point of the coupled walks lies in the sensor
                # ball, so it must sit inside the domain.
                if self._clearance(name, z) > SENSOR_RADII * r + 1e-3:
                    self.requests.append(Request(i, t, "validation", name, z[None, :], int(rng.integers(1 << 62)),
                                                 0.0, extra={"center": z, "radius": r}))
                    continue
            # A TD query (also the fallback when no probe lies inside).
            pick = rng.choice(len(w.sites), size=min(count, len(w.sites)), replace=count > len(w.sites))
            sites = w.sites[pick]
            self.requests.append(Request(i, t, "td", name, sites, int(rng.integers(1 << 62)),
                                         float(rng.uniform(ci_lo, ci_hi))))
            scored[name].append(sites[0])

    def _clearance(self, name: str, z: np.ndarray) -> float:
        """Distance from z to the base domain's boundary (0 outside a mesh)."""
        w, g = self.geoms[name]
        d, _, _ = self.ctx.closest(g, z[None, :], precision=64)
        if w.container == self.twos.MESH and not self.ctx.inside(self.mesh, z[None, :])[0]:
            return 0.0
        return float(d[0])

    def derived_geometry(self, req: Request) -> int:
        """The validation's geometry: the base domain with the candidate
        sphere's radius as its cavity radius (COUPLED cuts the sphere at the
        request's site for the second path of each walk)."""
        w, _ = self.geoms[req.geom_name]
        t, r = self.twos, req.extra["radius"]
        if w.container == t.MESH:
            return self.ctx.geometry(t.MESH, mesh=self.mesh, cavity_radius=r)
        if w.container == t.BOX:
            return self.ctx.geometry(t.BOX, box_lo=w.box_lo, box_hi=w.box_hi, spheres=w.spheres,
                                     cylinders=w.cylinders, cavity_radius=r)
        return self.ctx.geometry(t.BALL, ball=w.ball, cavity_radius=r)

    def request_args(self, req: Request, geom: int, distance_grid: bool = True):
        """(geometry, sites, sides, keyword arguments) of the request's
        estimate; the server and a standalone twos.Context.estimate take the
        same ones."""
        w, _ = self.geoms[req.geom_name]
        sides = W.sides(w, self.twos.side)
        common = dict(precision=32, seed=req.seed, eps=w.eps, max_steps=w.max_steps, distance_grid=distance_grid,
                      **self.shapes.get(req.geom_name, {}))
        if req.kind == "td":
            return geom, req.sites, sides, dict(combine=self.twos.CROSS, walks=self.walks, walks_max=self.walks_max,
                                                ci_rel=req.ci_rel, **common)
        probe = self.twos.probe_side(req.extra["center"], SENSOR_RADII * req.extra["radius"])
        return geom, req.sites, [sides[0], probe], dict(combine=self.twos.COUPLED, walks=VALIDATION_WALKS, **common)

    def submit(self, srv, req: Request, distance_grid: bool = True) -> None:
        _, g = self.geoms[req.geom_name]
        if req.kind == "validation":
            g = self.derived_geometry(req)
            req.extra["derived"] = g
        geom, sites, sides, kw = self.request_args(req, g, distance_grid)
        req.ids = [srv.submit(geom, sites, sides, **kw)]


def run(stream: Stream, srv, time_scale: float = 1.0, distance_grid: bool = True,
        log=None) -> dict:
    """Replays the stream in real time against the server; returns latencies
    and throughput."""
    pending = list(stream.requests)
    by_id: dict[int, Request] = {}
    inflight: dict[int, Request] = {}
    t0 = time.perf_counter()
    steps_total = 0
    sites_at_target = 0
    sites_total = 0
    rounds = 0
    failed = 0
    busy_s = 0.0                      # host time inside server rounds
    mesh_queries = 0                  # closest-point queries of mesh requests (steps + paths)
    while pending or inflight:
        now = time.perf_counter() - t0
        while pending and pending[0].arrival * time_scale <= now:
            req = pending.pop(0)
            stream.submit(srv, req, distance_grid)
            for i in req.ids:
                by_id[i] = req
            inflight[req.rid] = req
        if not inflight:
            time.sleep(max(0.0, pending[0].arrival * time_scale - now))
            continue
        t_step = time.perf_counter()
        done = srv.step()
        busy_s += time.perf_counter() - t_step
        rounds += 1
        t_end = time.perf_counter() - t0
        for sid, res in done.items():
            req = by_id.pop(sid)
            req.results.append((sid, res))
            req.done += 1
            steps_total += int(res.steps.sum())
            failed += res.stats.get("status", 0) != 0
            if stream.geoms[req.geom_name][0].container == stream.twos.MESH:
                mesh_queries += int(res.steps.sum()) + 2 * int(res.walks.sum())
            if req.done == len(req.ids):
                req.finish = t_end
                inflight.pop(req.rid)
                if req.kind == "td":
                    ok = (res.status == 0) & (1.96 * res.score_se <= req.ci_rel * np.abs(res.score))
                    sites_at_target += int(ok.sum())
                    sites_total += len(res.score)
                else:
                    stream.ctx.release_geometry(req.extra["derived"])
        if log and rounds % 50 == 0:
            log(f"round {rounds}: t={t_end:.2f}s inflight={len(inflight)} pending={len(pending)}")
    wall = time.perf_counter() - t0
    lat = np.array([r.finish - r.arrival * time_scale for r in stream.requests])
    return {"requests": len(stream.requests), "rounds": rounds, "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "td_sites_at_target": sites_at_target,
            "td_sites": sites_total, "td_sites_per_s": sites_at_target / wall, "walk_steps_per_s": steps_total / wall,
            "walk_steps": steps_total, "failed": failed, "busy_s": busy_s, "mesh_queries": mesh_queries,
            "validations": sum(r.kind == "validation" for r in stream.requests)}


def validation_change(req: Request) -> tuple[float, float]:
    """A validation's objective change per capacity,
    (J(domain) - J(domain minus the sphere)) / (4 pi r), and its standard
    error."""
    (_, res), = req.results
    return float(res.score[0]), float(res.score_se[0])
