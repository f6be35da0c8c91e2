"""Request stream over the continuous-batching server (configuration 4).

Resident geometries: the unit ball, the cooling block and the SH icosphere.
Requests arrive by a Poisson process; each has a log-uniform number of sites
in [1, max_sites], a CI target U[ci_lo, ci_hi] of |TD| and its own seed.  A
fraction are fresh-seed insertion validations: a sphere of radius U[0.01,
0.05] is unioned into the geometry at a previously scored site, and the
change of the primal state at probe points around it is estimated with
common streams (the same seed and walk indices on the geometry with and
without the cavity, so paths coincide until they feel the cavity).

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

VALIDATION_WALKS = 1024


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
                # Probes on a shell around the inserted sphere.
                k = max(1, min(count, 256))
                d = rng.normal(size=(k, 3))
                d /= np.linalg.norm(d, axis=1)[:, None]
                probes = z + d * (r + rng.uniform(0.5, 2.0, (k, 1)) * r)
                probes = probes[self._interior(name, probes)]
                if len(probes):
                    self.requests.append(Request(i, t, "validation", name, probes, int(rng.integers(1 << 62)),
                                                 0.0, extra={"center": z, "radius": r}))
                    continue
            # A TD query (also the fallback when no probe lies inside).
            pick = rng.choice(len(w.sites), size=min(count, len(w.sites)), replace=count > len(w.sites))
            sites = w.sites[pick]
            self.requests.append(Request(i, t, "td", name, sites, int(rng.integers(1 << 62)),
                                         float(rng.uniform(ci_lo, ci_hi))))
            scored[name].append(sites[0])

    def _interior(self, name: str, pts: np.ndarray) -> np.ndarray:
        """Probes strictly inside the base domain (a walk must start inside)."""
        w, g = self.geoms[name]
        d, _, _ = self.ctx.closest(g, pts, precision=64)
        keep = d > 1e-3
        if w.container == self.twos.MESH:
            keep &= self.ctx.inside(self.mesh, pts)
        return keep

    def submit(self, srv, req: Request, distance_grid: bool = True) -> None:
        w, g = self.geoms[req.geom_name]
        sides = W.sides(w, self.twos.side)
        common = dict(combine=self.twos.CROSS, precision=32, seed=req.seed, eps=w.eps, max_steps=w.max_steps,
                      distance_grid=distance_grid, **self.shapes.get(req.geom_name, {}))
        if req.kind == "td":
            req.ids = [srv.submit(g, req.sites, sides, walks=self.walks, walks_max=self.walks_max,
                                  ci_rel=req.ci_rel, **common)]
            return
        # Validation: the primal state at the same probes with the same
        # streams, with and without the cavity (SQUARE keeps sum a^2 for the SE).
        g2 = self.ctx.union_sphere(g, req.extra["center"], req.extra["radius"])
        req.extra["derived"] = g2
        common["combine"] = self.twos.SQUARE
        req.ids = [srv.submit(g, req.sites, sides[:1], walks=VALIDATION_WALKS, **common),
                   srv.submit(g2, req.sites, sides[:1], walks=VALIDATION_WALKS, **common)]


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
        done = srv.step()
        rounds += 1
        t_end = time.perf_counter() - t0
        for sid, res in done.items():
            req = by_id.pop(sid)
            req.results.append((sid, res))
            req.done += 1
            steps_total += int(res.steps.sum())
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
            "validations": sum(r.kind == "validation" for r in stream.requests)}


def validation_change(req: Request) -> tuple[np.ndarray, np.ndarray]:
    """Change of the primal state at the probes (without minus with the
    cavity) and a conservative standard error (the two estimates share their
    streams, which only lowers the true error)."""
    (_, base), (_, cav) = sorted(req.results, key=lambda x: x[0])
    se = lambda r: np.sqrt(np.maximum(r.same - r.u * r.u, 0.0) / r.walks)
    return base.u - cav.u, np.hypot(se(base), se(cav))
