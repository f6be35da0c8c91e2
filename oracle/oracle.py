"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and there only as the checker.  The product library never does.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

BALL, BOX, MESH = 0, 1, 2
SQUARE, PAIRED, CROSS, COUPLED = 0, 1, 2, 3
RNG_PHILOX, RNG_MT_SITE, RNG_MT_PAIRED, RNG_MT_KEY, RNG_MT_FD, RNG_MT_DIFF = 0, 1, 2, 3, 4, 5

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Geom(C.Structure):
    _fields_ = [("container", C.c_int32), ("n_spheres", C.c_int32), ("n_cylinders", C.c_int32),
                ("n_cavities", C.c_int32), ("mesh_radius_f32", C.c_int32), ("pad", C.c_int32),
                ("ball", C.c_double * 4), ("box_lo", C.c_double * 3),
                ("box_hi", C.c_double * 3), ("cavity_radius", C.c_double), ("spheres", _dp),
                ("cylinders", _dp), ("cavities", _dp), ("mesh", C.c_void_p)]


MAX_TERMS = 8
TERM_BUMP, TERM_UNIT, TERM_GAUSS = 0, 1, 2


class Side(C.Structure):
    _fields_ = [("poly", C.c_double * 10), ("cls_const", C.c_double * 8),
                ("cls_poly_mask", C.c_uint32), ("source_kind", C.c_int32), ("src", C.c_double * 7),
                ("n_terms", C.c_int32), ("pad", C.c_int32), ("terms", (C.c_double * 8) * MAX_TERMS)]


class Params(C.Structure):
    _fields_ = [("combine", C.c_int32), ("rng", C.c_int32), ("seed", C.c_uint64),
                ("repeat", C.c_int32), ("max_steps", C.c_int32), ("eps", C.c_double),
                ("walks", C.c_int32), ("walks_max", C.c_int32), ("ci_rel", C.c_double),
                ("ci_z", C.c_double), ("threads", C.c_int32), ("gradient", C.c_int32)]


class Site(C.Structure):
    _fields_ = [("u", C.c_double), ("v", C.c_double), ("score", C.c_double),
                ("score_se", C.c_double), ("td", C.c_double), ("same", C.c_double),
                ("walks", C.c_int64), ("steps", C.c_int64), ("caps", C.c_int64),
                ("status", C.c_int32), ("pad", C.c_int32), ("grad_u", C.c_double * 3),
                ("grad_v", C.c_double * 3), ("grad_dot", C.c_double), ("grad_dot_se", C.c_double)]


def build() -> Path:
    """Compile the oracle with its Makefile (g++, OpenMP, binary64)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.og_splitmix64.restype = C.c_uint64
        L.og_splitmix64.argtypes = [C.c_uint64]
        L.og_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.og_philox_key.argtypes = [C.c_uint64, C.c_int64, C.c_int32, _u32p]
        L.og_mt19937_64.argtypes = [C.c_uint64, C.POINTER(C.c_uint64), C.c_int32]
        L.og_mt_uniforms.argtypes = [C.c_uint64, _dp, C.c_int32]
        L.og_mesh_create.restype = C.c_void_p
        L.og_mesh_create.argtypes = [_dp, C.c_int64, _u32p, C.c_int64]
        L.og_mesh_destroy.argtypes = [C.c_void_p]
        L.og_mesh_closest.restype = C.c_double
        L.og_mesh_closest.argtypes = [C.c_void_p, _dp, _dp, _i64p, C.c_int32, _i64p, _i64p]
        L.og_closest_point_triangle.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.og_mesh_inside.restype = C.c_int32
        L.og_mesh_inside.argtypes = [C.c_void_p, _dp]
        L.og_mesh_inside_many.restype = C.c_int32
        L.og_mesh_inside_many.argtypes = [C.c_void_p, _dp, C.c_int64, _i32p, C.c_int32]
        L.og_closest_boundary.restype = C.c_double
        L.og_closest_boundary.argtypes = [C.POINTER(Geom), _dp, _dp, _i32p]
        L.og_estimate.restype = C.c_int32
        L.og_estimate.argtypes = [C.POINTER(Geom), C.POINTER(Side), C.POINTER(Params), _dp,
                                  C.c_int64, C.POINTER(Site), _dp, _dp, _i32p, C.c_int64]
        L.og_trace.restype = C.c_int32
        L.og_trace.argtypes = [C.POINTER(Geom), C.POINTER(Side), C.POINTER(Params), _dp,
                               C.c_uint64, C.c_int64, C.c_int32, _dp, C.c_int32]
        L.og_combine.argtypes = [_dp, _dp, C.c_int64, C.c_int32, _dp]
        L.og_combine_powersums.argtypes = [_dp, C.c_int64, _dp]
        L.og_combine_grad.argtypes = [_dp, _dp, C.c_int64, _dp]
        L.og_bump_green_ball.argtypes = [_dp, C.c_double, _dp, C.c_double, C.c_double]
        L.og_differential.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, C.c_int64, C.c_double, _dp,
                                      C.POINTER(C.c_int64)]
        L.og_differential.restype = C.c_int32
        L.og_bump_green_ball.restype = C.c_double
        L.og_certify.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_int64, _dp, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)]
        L.og_certify.restype = C.c_int32
        L.og_cert_distance.argtypes = [C.c_void_p, _dp, C.c_int64, _dp]
        L.og_cert_distance.restype = C.c_int32
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def splitmix64(x: int) -> int:
    return int(lib().og_splitmix64(x))


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().og_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
    return out


def philox_key(seed: int, site: int, side: int) -> np.ndarray:
    out = np.zeros(2, dtype=np.uint32)
    lib().og_philox_key(seed, site, side, _ptr(out, _u32p))
    return out


def mt19937_64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().og_mt19937_64(seed, out.ctypes.data_as(C.POINTER(C.c_uint64)), n)
    return out


def mt_uniforms(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().og_mt_uniforms(seed, _ptr(out, _dp), n)
    return out


class Mesh:
    def __init__(self, vertices, faces):
        self.vertices = _f64(vertices)
        self.faces = np.ascontiguousarray(faces, dtype=np.uint32)
        self.handle = lib().og_mesh_create(_ptr(self.vertices, _dp), len(self.vertices),
                                           _ptr(self.faces, _u32p), len(self.faces))
        if not self.handle:
            raise ValueError("invalid mesh")

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.og_mesh_destroy(self.handle)
            self.handle = None

    def closest(self, p, brute: bool = False):
        p = _f64(p)
        q = np.zeros(3)
        tri = C.c_int64()
        nv, nt = C.c_int64(0), C.c_int64(0)
        d = lib().og_mesh_closest(self.handle, _ptr(p, _dp), _ptr(q, _dp), C.byref(tri),
                                  int(brute), C.byref(nv), C.byref(nt))
        return d, q, tri.value, nv.value, nt.value

    def inside(self, points, threads: int = 0) -> np.ndarray:
        p = _f64(np.atleast_2d(points))
        out = np.zeros(len(p), dtype=np.int32)
        lib().og_mesh_inside_many(self.handle, _ptr(p, _dp), len(p), _ptr(out, _i32p), threads)
        return out


@dataclass
class Geometry:
    container: int
    ball: tuple = (0.0, 0.0, 0.0, 1.0)
    box_lo: tuple = (0.0, 0.0, 0.0)
    box_hi: tuple = (1.0, 1.0, 1.0)
    spheres: np.ndarray | None = None
    cylinders: np.ndarray | None = None
    cavities: np.ndarray | None = None
    cavity_radius: float = 0.0
    mesh: Mesh | None = None
    mesh_radius_f32: bool = False

    def struct(self) -> Geom:
        g = Geom()
        g.container = self.container
        g.mesh_radius_f32 = int(self.mesh_radius_f32)
        g.ball = (C.c_double * 4)(*self.ball)
        g.box_lo = (C.c_double * 3)(*self.box_lo)
        g.box_hi = (C.c_double * 3)(*self.box_hi)
        self._keep = []
        for name, count in (("spheres", "n_spheres"), ("cylinders", "n_cylinders"),
                            ("cavities", "n_cavities")):
            arr = getattr(self, name)
            if arr is not None and len(arr):
                a = _f64(arr)
                self._keep.append(a)
                setattr(g, name, _ptr(a, _dp))
                setattr(g, count, len(a))
        g.cavity_radius = self.cavity_radius
        g.mesh = self.mesh.handle if self.mesh is not None else None
        return g


def side(poly=None, cls_const=None, poly_classes=(), source=None, terms=None, offset=0.0,
         analytic: bool = False) -> Side:
    """Boundary data g = cls_const[class] + [class in poly_classes] * poly(y);
    source = (s0, sx, sy, sz, ox, oy, oz) adds (r^2/6) f(x) per step at the ball
    centre; terms = rows (type, cx, cy, cz, p0, p1, p2, w) add (r^2/6) f(a) with
    f = offset + sum of terms at a Green-distributed point a, or with
    analytic=True the exact Green-ball integral of offset and each BUMP / UNIT
    row."""
    s = Side()
    if terms is not None:
        if source is not None:
            raise ValueError("a side has either an affine source or sampled terms")
        terms = [list(map(float, t)) for t in terms]
        if not 0 < len(terms) <= MAX_TERMS or any(len(t) != 8 for t in terms):
            raise ValueError(f"expected 1..{MAX_TERMS} term rows of 8 numbers")
        s.source_kind = 3 if analytic else 2
        s.src[0] = offset
        s.n_terms = len(terms)
        for k, t in enumerate(terms):
            s.terms[k] = (C.c_double * 8)(*t)
    s.poly = (C.c_double * 10)(*(list(poly) if poly is not None else [0.0] * 10))
    s.cls_const = (C.c_double * 8)(*(list(cls_const) if cls_const is not None else [0.0] * 8))
    mask = 0
    for c in poly_classes:
        mask |= 1 << c
    s.cls_poly_mask = mask
    if source is not None:
        s.source_kind = 1
        s.src = (C.c_double * 7)(*source)
    return s


def reference_affine_side(slope: float = 0.65) -> Side:
    """f = 1 + slope * (y + 0.5), zero terminal value (wih_twos.cpp:546-552)."""
    return side(source=(1.0, 0.0, slope, 0.0, 0.0, 0.5, 0.0))


def study_terms(config_text: str):
    """Heat and sensor source rows of a study configuration: first line the
    sensor (cx cy cz R), then heat bumps (cx cy cz R A) (study.cpp:37-43).
    The sensor's mass 32 pi R^3 / 105 is computed as study.cpp:50 does."""
    vals = [float(v) for v in config_text.split()]
    sensor, rest = vals[:4], vals[4:]
    heat = [rest[i:i + 5] for i in range(0, len(rest) - len(rest) % 5, 5)]
    if len(sensor) < 4 or not heat:
        raise ValueError("expected a sensor line and at least one heat bump")
    heat_rows = [(TERM_BUMP, h[0], h[1], h[2], h[3], 0.0, 0.0, h[4]) for h in heat]
    mass = 32 * 3.141592653589793 * float(np.power(sensor[3], 3.0)) / 105
    sensor_rows = [(TERM_UNIT, sensor[0], sensor[1], sensor[2], sensor[3], mass, 0.0, 0.0)]
    return heat_rows, sensor_rows


def housing_chip_terms():
    """The built-in housing source 0.08 + 5.2 chip_a + 4.6 chip_b + 2.0 regulator
    (wih_twos.cpp:553-566): (offset, rows)."""
    return 0.08, [(TERM_GAUSS, -0.48, -0.18, -0.08, 0.24, 0.22, 0.16, 5.2),
                  (TERM_GAUSS, 0.43, 0.19, -0.07, 0.21, 0.20, 0.15, 4.6),
                  (TERM_GAUSS, 0.02, -0.31, -0.02, 0.30, 0.16, 0.19, 2.0)]


def estimate(geometry: Geometry, sites, sides, *, combine=SQUARE, rng=RNG_PHILOX, seed=0,
             repeat=0, walks=16, walks_max=0, ci_rel=0.02, ci_z=1.96, eps=5e-4, max_steps=512,
             threads=0, gradient: bool = False, dump: bool = False):
    sites = _f64(np.atleast_2d(sites))
    n = len(sites)
    g = geometry.struct()
    sd = (Side * 2)(sides[0], sides[1] if len(sides) > 1 else sides[0])
    p = Params(combine=combine, rng=rng, seed=seed, repeat=repeat, max_steps=max_steps, eps=eps,
               walks=walks, walks_max=walks_max, ci_rel=ci_rel, ci_z=ci_z, threads=threads,
               gradient=int(gradient))
    out = (Site * max(n, 1))()
    stride = max(walks, walks_max) if dump else 0
    da = np.zeros((n, stride))
    db = np.zeros((n, stride))
    ds = np.zeros((n, stride), dtype=np.int32)
    status = lib().og_estimate(C.byref(g), sd, C.byref(p), _ptr(sites, _dp), n, out,
                               _ptr(da, _dp) if dump else None, _ptr(db, _dp) if dump else None,
                               _ptr(ds, _i32p) if dump else None, stride)
    if status == 2:
        raise ValueError("invalid estimator arguments")
    fields = [f[0] for f in Site._fields_ if f[0] != "pad"]
    res = {k: np.array([np.array(getattr(out[i], k)) for i in range(n)]) for k in fields}
    res["status_code"] = status
    if dump:
        res["a"], res["b"], res["walk_steps"] = da, db, ds
    return res


def trace(geometry: Geometry, start, side_, *, rng=RNG_MT_KEY, key=0, seed=0, site=0, walk=0,
          repeat=0, eps=5e-4, max_steps=512) -> np.ndarray:
    g = geometry.struct()
    p = Params(combine=SQUARE, rng=rng, seed=seed, repeat=repeat, max_steps=max_steps, eps=eps,
               walks=2)
    rows = np.zeros((max_steps + 2, 4))
    s = _f64(start)
    count = lib().og_trace(C.byref(g), C.byref(side_), C.byref(p), _ptr(s, _dp), key, site, walk,
                           _ptr(rows, _dp), len(rows))
    return rows[:count]


def combine(a, b=None, square: bool = False) -> np.ndarray:
    a = _f64(a)
    b = a if b is None else _f64(b)
    out = np.zeros(6)
    lib().og_combine(_ptr(a, _dp), _ptr(b, _dp), len(a), int(square), _ptr(out, _dp))
    return out


def combine_powersums(sums, n: int) -> np.ndarray:
    s = _f64(sums)
    out = np.zeros(6)
    lib().og_combine_powersums(_ptr(s, _dp), n, _ptr(out, _dp))
    return out


DIFF_COUNTS = ("outer_steps", "nested_steps", "cavity_terminations", "near_seam", "segment_unsafe", "unsafe",
               "limits", "distance_queries", "parity_rays")


def differential(geometry: Geometry, sites, side_, *, seed=0, repeat=0, walks=16, eps=1e-4, max_steps=2048,
                 offset=0.004, threads=0, rng=RNG_MT_DIFF):
    """Differential WoS cavity derivatives per site on the reference's
    streams (rng RNG_MT_DIFF) or on Philox streams (RNG_PHILOX): (values
    [n][n_cavities][8] = dc x y z, their SEs, dR, its SE; counts [n][9] in
    DIFF_COUNTS order)."""
    if rng not in (RNG_MT_DIFF, RNG_PHILOX):
        raise ValueError("differential streams are RNG_MT_DIFF or RNG_PHILOX")
    sites = _f64(np.atleast_2d(sites))
    n = len(sites)
    g = geometry.struct()
    k = g.n_cavities
    p = Params(rng=rng, seed=seed, repeat=repeat, max_steps=max_steps, eps=eps, walks=walks,
               threads=threads)
    out = np.zeros((n, max(k, 1), 8))
    counts = np.zeros((n, 9), dtype=np.int64)
    st = lib().og_differential(C.byref(g), C.byref(side_), C.byref(p), _ptr(sites, _dp), n, offset,
                               _ptr(out, _dp), counts.ctypes.data_as(C.POINTER(C.c_int64)))
    if st == 2:
        raise ValueError("invalid differential arguments")
    return out[:, :k], counts


class CertParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("walks", C.c_int32), ("max_steps", C.c_int32), ("eps", C.c_double),
                ("radius_cap", C.c_double), ("value_cap", C.c_double), ("slope", C.c_double),
                ("source", C.c_int32), ("threads", C.c_int32), ("site_offset", C.c_int64), ("rng", C.c_int32),
                ("pad", C.c_int32)]


CERT_AFFINE, CERT_HOUSING = 0, 1


def certify(geometry: Geometry, sites, *, seed=0, walks=1024, max_steps=128, eps=1e-6, radius_cap=0.1,
            value_cap=0.1, source=CERT_AFFINE, slope=0.65, threads=0, site_offset=0, rng=RNG_MT_KEY):
    """Certified minimum-benefit samples (cert_walk, minimum_benefit.hpp:122-143)
    on the reference's streams (rng RNG_MT_KEY) or Philox streams
    (RNG_PHILOX): (value, steps, flags) as [n][walks] arrays, flags 1 capped,
    2 clipped, 4 containment failure."""
    sites = _f64(np.atleast_2d(sites))
    n = len(sites)
    g = geometry.struct()
    p = CertParams(seed=seed, walks=walks, max_steps=max_steps, eps=eps, radius_cap=radius_cap,
                   value_cap=value_cap, slope=slope, source=source, threads=threads, site_offset=site_offset,
                   rng=rng)
    value = np.zeros((n, walks))
    steps = np.zeros((n, walks), dtype=np.int32)
    flags = np.zeros((n, walks), dtype=np.int32)
    st = lib().og_certify(C.byref(g), C.byref(p), _ptr(sites, _dp), n, _ptr(value, _dp),
                          steps.ctypes.data_as(C.POINTER(C.c_int32)), flags.ctypes.data_as(C.POINTER(C.c_int32)))
    if st == 2:
        raise ValueError("invalid certification arguments")
    return value, steps, flags


def cert_distance(geometry: Geometry, points) -> np.ndarray:
    """Certified lower bound of the domain distance (cert_domain_distance)."""
    pts = _f64(np.atleast_2d(points))
    out = np.zeros(len(pts))
    g = geometry.struct()
    if lib().og_cert_distance(C.byref(g), _ptr(pts, _dp), len(pts), _ptr(out, _dp)) == 2:
        raise ValueError("invalid geometry")
    return out


def bump_green_ball(x, radius: float, center, support: float, amplitude: float) -> float:
    """Exact integral of the Green function of B(x, radius) against the bump
    amplitude (1 - |y - center|^2 / support^2)^2."""
    return lib().og_bump_green_ball(_ptr(_f64(x), _dp), radius, _ptr(_f64(center), _dp), support, amplitude)


def combine_grad(ga, gb) -> np.ndarray:
    """Gradient channel from per-walk samples ([n][3] each): grad_u[3],
    grad_v[3], grad_dot, grad_dot_se."""
    a = _f64(np.asarray(ga).reshape(-1, 3))
    b = _f64(np.asarray(gb).reshape(-1, 3))
    out = np.zeros(8)
    lib().og_combine_grad(_ptr(a, _dp), _ptr(b, _dp), len(a), _ptr(out, _dp))
    return out


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
