"""Synthetic workloads of the estimator (all generators seeded and frozen).

unit_ball       Laplace in B(0,1), g_u = x^2 - y^2 + 0.5 z, g_v = x z + y; 64 sites
                uniform in the radius-0.8 ball; binary64; eps 1e-3.
cooling_block   [0,1]^3 minus 12 non-overlapping spheres (r ~ U[0.04, 0.10]) and two
                axis-aligned cylinders (r 0.05); u = 1 on holes, 0 on the faces; v = 1 on
                the z = 1 face; sites on the 16^3 cell-centre grid at distance >= 0.02.
sh_icosphere    level-6 icosphere (81,920 triangles) displaced radially by a seeded
                real spherical-harmonic field (l <= 8, amplitude 0.15), scaled into the unit
                box; g_u = x + 0.3 y^2, g_v = z; 16,384 interior sites by rejection.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    sites: np.ndarray
    poly_u: list
    poly_v: list
    eps: float
    precision: int
    container: int
    cls_u: list = field(default_factory=lambda: [0.0] * 8)
    cls_v: list = field(default_factory=lambda: [0.0] * 8)
    poly_classes: tuple = (0,)
    ball: tuple = (0.0, 0.0, 0.0, 1.0)
    box_lo: tuple = (0.0, 0.0, 0.0)
    box_hi: tuple = (1.0, 1.0, 1.0)
    spheres: np.ndarray | None = None
    cylinders: np.ndarray | None = None
    vertices: np.ndarray | None = None
    faces: np.ndarray | None = None
    walks: int = 1024
    walks_max: int = 0
    ci_rel: float = 0.02
    max_steps: int = 8192


def _ball_sites(n: int, radius: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d * (radius * rng.uniform(0.0, 1.0, n) ** (1.0 / 3.0))[:, None]


def unit_ball(n_sites: int = 64, walks: int = 1024) -> Workload:
    return Workload(name="unit_ball_harmonic", sites=_ball_sites(n_sites, 0.8, 0),
                    poly_u=[0, 0, 0, 0.5, 1, -1, 0, 0, 0, 0], poly_v=[0, 0, 1, 0, 0, 0, 0, 0, 1, 0],
                    eps=1e-3, precision=64, container=0, walks=walks)


def unit_ball_exact(sites: np.ndarray) -> np.ndarray:
    """D_T = -g_u g_v at the sites (u = g_u, v = g_v for harmonic data)."""
    x, y, z = sites.T
    return -(x * x - y * y + 0.5 * z) * (x * z + y)


def _place_spheres(count: int, seed: int, clearance: float = 0.02):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        r = rng.uniform(0.04, 0.10)
        c = rng.uniform(r + clearance, 1.0 - r - clearance, 3)
        if all(np.linalg.norm(c - o[:3]) > r + o[3] + clearance for o in out):
            out.append(np.array([*c, r]))
    return np.array(out)


def _box_distance(p, spheres, cylinders):
    d = np.minimum(p, 1.0 - p).min(axis=1)
    for s in spheres:
        d = np.minimum(d, np.linalg.norm(p - s[:3], axis=1) - s[3])
    for c in cylinders:
        axis = int(c[0])
        a, b = [k for k in range(3) if k != axis]
        d = np.minimum(d, np.hypot(p[:, a] - c[1], p[:, b] - c[2]) - c[3])
    return d


def cooling_block(walks: int = 8192) -> Workload:
    spheres = _place_spheres(12, 2)
    rng = np.random.default_rng(3)
    cylinders = []
    for axis in rng.choice(3, size=2, replace=True):
        cylinders.append([float(axis), *rng.uniform(0.15, 0.85, 2), 0.05])
    cylinders = np.array(cylinders)
    g = (np.arange(16) + 0.5) / 16.0
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    keep = _box_distance(grid, spheres, cylinders) >= 0.02
    cls_u = [0.0] * 6 + [1.0, 0.0]
    cls_v = [0.0] * 5 + [1.0, 0.0, 0.0]
    return Workload(name="cooling_block", sites=grid[keep], poly_u=[0.0] * 10, poly_v=[0.0] * 10, eps=1e-4,
                    precision=32, container=1, cls_u=cls_u, cls_v=cls_v, poly_classes=(), spheres=spheres,
                    cylinders=cylinders, walks=walks)


def icosphere(level: int):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = np.array(f, dtype=np.int64)
    for _ in range(level):
        cache: dict = {}
        new = []

        def mid(i, j):
            key = (i, j) if i < j else (j, i)
            k = cache.get(key)
            if k is None:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                k = cache[key] = len(verts) - 1
            return k

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = np.array(new, dtype=np.int64)
    return np.array(verts), faces.astype(np.uint32)


def _legendre_table(lmax: int, x: np.ndarray) -> dict:
    """Associated Legendre P_l^m(x), m >= 0, with the Condon-Shortley phase,
    by the standard three-term recurrences."""
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    p = {(0, 0): np.ones_like(x)}
    for m in range(1, lmax + 1):
        p[(m, m)] = -(2 * m - 1) * s * p[(m - 1, m - 1)]
    for m in range(0, lmax):
        p[(m + 1, m)] = (2 * m + 1) * x * p[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            p[(l, m)] = ((2 * l - 1) * x * p[(l - 1, m)] - (l + m - 1) * p[(l - 2, m)]) / (l - m)
    return p


def sh_field(directions: np.ndarray, lmax: int, seed: int) -> np.ndarray:
    """F(w) = sum_{l<=lmax, |m|<=l} c_lm Y_lm(w), real orthonormal Y_lm, c_lm ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    x, y, z = directions.T
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    p = _legendre_table(lmax, np.cos(theta))
    f = np.zeros(len(directions))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            if m == 0:
                y_lm = norm * p[(l, 0)]
            elif m > 0:
                y_lm = math.sqrt(2.0) * norm * p[(l, am)] * np.cos(am * phi)
            else:
                y_lm = math.sqrt(2.0) * norm * p[(l, am)] * np.sin(am * phi)
            f += rng.normal() * y_lm
    return f


def sh_icosphere_mesh(level: int = 6, amplitude: float = 0.15, lmax: int = 8, seed: int = 5):
    """Returns vertices, faces and the map of the smooth surface:
    r(w) = scale * (1 + amplitude F(w) / fmax) about `centre`."""
    v, f = icosphere(level)
    field_ = sh_field(v, lmax, seed)
    fmax = float(np.abs(field_).max())
    v = v * (1.0 + amplitude * field_ / fmax)[:, None]
    lo, hi = v.min(0), v.max(0)
    scale = 1.0 / (hi - lo).max()
    v = (v - 0.5 * (lo + hi)) * scale + 0.5
    centre = 0.5 - 0.5 * (lo + hi) * scale
    return v, f, {"centre": centre, "scale": scale, "amplitude": amplitude, "lmax": lmax, "seed": seed,
                  "fmax": fmax}


def sh_icosphere(n_sites: int = 16384, margin: float = 0.01, seed: int = 7) -> Workload:
    """Sites: uniform in the mesh's bounding box, kept when their radius from
    the centre is below the smooth displaced radius minus `margin` (the level-6
    chords sag by < 1e-3 below the smooth surface, so every kept site lies
    inside the mesh at distance > 0.009)."""
    v, f, s = sh_icosphere_mesh()
    rng = np.random.default_rng(seed)
    lo, hi = v.min(0), v.max(0)
    out = []
    got = 0
    while got < n_sites:
        p = rng.uniform(lo, hi, (4 * n_sites, 3))
        d = p - s["centre"]
        rad = np.linalg.norm(d, axis=1)
        fw = sh_field(d / rad[:, None], s["lmax"], s["seed"])
        r_smooth = (1.0 + s["amplitude"] * fw / s["fmax"]) * s["scale"]
        keep = p[rad < r_smooth - margin]
        out.append(keep)
        got += len(keep)
    sites = np.concatenate(out)[:n_sites]
    return Workload(name="sh_icosphere_81920", sites=sites, poly_u=[0, 1, 0, 0, 0, 0.3, 0, 0, 0, 0],
                    poly_v=[0, 0, 0, 1, 0, 0, 0, 0, 0, 0], eps=1e-4, precision=32, container=2, vertices=v,
                    faces=f, walks=256, walks_max=16384, ci_rel=0.02)


def sides(w: Workload, side_fn):
    """(primal, adjoint) boundary data of a workload built with `side_fn`
    (twos.side or the oracle's side)."""
    return [side_fn(poly=w.poly_u, cls_const=w.cls_u, poly_classes=w.poly_classes),
            side_fn(poly=w.poly_v, cls_const=w.cls_v, poly_classes=w.poly_classes)]
