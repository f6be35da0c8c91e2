"""MI355X Topological Walk on Spheres (TWoS) estimator.

Python host over the gfx950 HIP library (include/twos.h via ctypes).  The
estimator call keeps the reference backend's meaning: geometry, candidate
sites, a walk budget and a seed in; per site the state mean u, the
cross-index score (Sa Sb - Sab)/(N(N-1)), its jackknife standard error and the
step counts out (benchmarks/compiled_bunny/src/wih_twos.cpp:674-717,
benchmarks/submission_study/study.cpp:80-120).  The topological derivative is
td = -score.

    ctx = twos.Context()
    mesh = ctx.mesh(vertices, faces)
    geom = ctx.geometry(twos.MESH, mesh=mesh)
    res = ctx.estimate(geom, sites, [primal, adjoint], combine=twos.CROSS,
                       walks=256, walks_max=16384, ci_rel=0.02)
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import (BALL, BOX, CERT_AFFINE, CERT_CAPPED, CERT_CLIPPED, CERT_CONTAINMENT, CERT_HOUSING, COUPLED, CROSS, MESH, PAIRED, PERSISTENT, PLAIN, RNG_MT_FD, RNG_MT_KEY,
                    RNG_MT_PAIRED, RNG_MT_SITE, RNG_PHILOX, SQUARE, FLAG_COUNT_TRAVERSAL, FLAG_DISTANCE_GRID, FLAG_GRADIENT, MAX_TERMS,
                    TERM_BUMP, TERM_GAUSS, TERM_UNIT)

__all__ = [
    "Context", "Mesh", "Side", "Result", "Server", "TwosError", "InvalidArgument", "StepCapError",
    "GeometryError", "AdmissionError", "DeviceError", "NotFoundError", "side", "reference_affine_side",
    "BALL", "BOX", "MESH", "SQUARE", "PAIRED", "CROSS", "COUPLED", "RNG_PHILOX", "RNG_MT_SITE", "RNG_MT_PAIRED",
    "RNG_MT_KEY", "RNG_MT_FD", "probe_side", "PLAIN", "PERSISTENT", "FLAG_COUNT_TRAVERSAL", "FLAG_DISTANCE_GRID", "FLAG_GRADIENT", "version",
    "TERM_BUMP", "TERM_UNIT", "TERM_GAUSS", "MAX_TERMS", "study_terms", "housing_chip_terms",
    "CERT_AFFINE", "CERT_HOUSING", "CERT_CAPPED", "CERT_CLIPPED", "CERT_CONTAINMENT", "RESULT_BYTES",
    "results_from_bytes",
]


class TwosError(RuntimeError):
    code = -1

    def __init__(self, message: str, code: int | None = None, result=None):
        super().__init__(message)
        if code is not None:
            self.code = code
        self.result = result


class InvalidArgument(TwosError, ValueError):
    code = _capi.ERR_INVALID_ARGUMENT


class StepCapError(TwosError):
    """One or more paths hit max_steps (the reference's fatal
    "one or more paths hit --max-steps", wih_twos.cpp:1846-1848).  The
    estimate is attached as `.result`."""
    code = _capi.ERR_STEP_CAP


class GeometryError(TwosError):
    code = _capi.ERR_GEOMETRY


class AdmissionError(TwosError):
    code = _capi.ERR_ADMISSION


class DeviceError(TwosError):
    code = _capi.ERR_DEVICE


class NotFoundError(TwosError, KeyError):
    code = _capi.ERR_NOT_FOUND


_ERRORS = {c.code: c for c in (InvalidArgument, StepCapError, GeometryError, AdmissionError, DeviceError,
                               NotFoundError)}


def _lib():
    return _capi.load()


def _check(status: int, result=None) -> None:
    if status != _capi.OK:
        msg = _lib().twos_last_error().decode(errors="replace")
        raise _ERRORS.get(status, TwosError)(msg, status, result)


def version() -> str:
    return _lib().twos_version().decode()


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


Side = _capi.Side


def side(poly=None, cls_const=None, poly_classes=(), source=None, terms=None, offset=0.0,
         analytic: bool = False) -> Side:
    """Boundary data and source of one side.  The Dirichlet value at the closest
    boundary point y of class c is cls_const[c] + [c in poly_classes] * poly(y),
    poly = (c0, cx, cy, cz, cxx, cyy, czz, cxy, cxz, cyz).  Classes: 0-5 container
    faces (ball and mesh use 0; box faces -x,+x,-y,+y,-z,+z), 6 holes, 7 cavities.
    source = (s0, sx, sy, sz, ox, oy, oz) adds (r^2/6) f(x) per step with
    f = ((s0 + sx (x+ox)) + sy (y+oy)) + sz (z+oz) at the ball centre.
    terms = rows (type, cx, cy, cz, p0, p1, p2, w) (TERM_BUMP, TERM_UNIT,
    TERM_GAUSS) add (r^2/6) f(a) with f = offset + sum of the rows, at a
    Green-distributed point a of each step's ball (wih_twos.cpp:648-653);
    with analytic=True each step adds instead the exact integral of the ball's
    Green function against offset and each BUMP / UNIT row (the reference's
    analytic-source variant, study.cpp:56-76), with no random draw."""
    s = Side()
    if terms is not None:
        if source is not None:
            raise InvalidArgument("a side has either an affine source or sampled terms")
        rows = [[float(v) for v in t] for t in terms]
        if not 0 < len(rows) <= MAX_TERMS or any(len(t) != 8 for t in rows):
            raise InvalidArgument(f"terms needs 1..{MAX_TERMS} rows of 8 numbers")
        if analytic and any(int(t[0]) == TERM_GAUSS for t in rows):
            raise InvalidArgument("the analytic source integrates BUMP and UNIT rows only")
        s.source_kind = 3 if analytic else 2
        s.src[0] = float(offset)
        s.n_terms = len(rows)
        for k, t in enumerate(rows):
            s.terms[k] = (C.c_double * 8)(*t)
    poly = list(poly) if poly is not None else [0.0] * 10
    if len(poly) != 10:
        raise InvalidArgument("poly needs 10 coefficients")
    s.poly = (C.c_double * 10)(*poly)
    cc = list(cls_const) if cls_const is not None else [0.0] * 8
    if len(cc) != 8:
        raise InvalidArgument("cls_const needs 8 entries")
    s.cls_const = (C.c_double * 8)(*cc)
    mask = 0
    for c in poly_classes:
        if not 0 <= c < 8:
            raise InvalidArgument("boundary class out of range")
        mask |= 1 << c
    s.cls_poly_mask = mask
    if source is not None:
        if len(source) != 7:
            raise InvalidArgument("source needs (s0, sx, sy, sz, ox, oy, oz)")
        s.source_kind = 1
        s.src = (C.c_double * 7)(*source)
    return s


def probe_side(center, radius: float) -> Side:
    """The objective density of a COUPLED estimate: the unit-mass bump
    (1 - |x - c|^2 / R^2)^2 / (32 pi R^3 / 105) on the ball B(c, R)
    (study.cpp:44-51); COUPLED draws each walk's start point uniformly in that
    ball with weight (1 - rho^2)^2 35/8."""
    r = float(radius)
    mass = 32 * 3.141592653589793 * float(np.power(r, 3.0)) / 105
    return side(terms=[(TERM_UNIT, *[float(c) for c in center], r, mass, 0.0, 0.0)])


def reference_affine_side(slope: float = 0.65) -> Side:
    """The reference's compliance problem: zero Dirichlet data, source
    f = 1 + slope * (y + 0.5) (wih_twos.cpp:546-552, --source-y-slope)."""
    return side(source=(1.0, 0.0, slope, 0.0, 0.0, 0.5, 0.0))


def study_terms(config_text: str):
    """(heat rows, sensor rows) of a study source configuration: the first
    line is the sensor (cx cy cz R), then heat bumps (cx cy cz R A)
    (benchmarks/submission_study/study.cpp:37-43).  The sensor's mass
    32 pi R^3 / 105 is computed as study.cpp:50 computes it."""
    vals = [float(v) for v in config_text.split()]
    if len(vals) < 9 or (len(vals) - 4) % 5:
        raise InvalidArgument("expected a sensor line and at least one heat bump")
    sensor, rest = vals[:4], vals[4:]
    heat = [(TERM_BUMP, *rest[i:i + 3], rest[i + 3], 0.0, 0.0, rest[i + 4]) for i in range(0, len(rest), 5)]
    if len(heat) > MAX_TERMS:
        raise InvalidArgument(f"at most {MAX_TERMS} heat bumps")
    mass = 32 * 3.141592653589793 * float(np.power(sensor[3], 3.0)) / 105
    return heat, [(TERM_UNIT, *sensor[:3], sensor[3], mass, 0.0, 0.0)]


def housing_chip_terms():
    """(offset, rows) of the built-in housing source
    0.08 + 5.2 chip_a + 4.6 chip_b + 2.0 regulator (wih_twos.cpp:553-566)."""
    return 0.08, [(TERM_GAUSS, -0.48, -0.18, -0.08, 0.24, 0.22, 0.16, 5.2),
                  (TERM_GAUSS, 0.43, 0.19, -0.07, 0.21, 0.20, 0.15, 4.6),
                  (TERM_GAUSS, 0.02, -0.31, -0.02, 0.30, 0.16, 0.19, 2.0)]


@dataclass
class Mesh:
    id: int
    n_vertices: int
    n_faces: int
    build_ms: float
    depth: int
    bvh_bytes: int


@dataclass
class Result:
    u: np.ndarray
    v: np.ndarray
    score: np.ndarray
    score_se: np.ndarray
    td: np.ndarray
    same: np.ndarray
    walks: np.ndarray
    steps: np.ndarray
    caps: np.ndarray
    status: np.ndarray
    stats: dict = field(default_factory=dict)
    grad_u: np.ndarray | None = None    # [n][3] with gradient=True
    grad_v: np.ndarray | None = None
    grad_dot: np.ndarray | None = None
    grad_dot_se: np.ndarray | None = None
    a: np.ndarray | None = None
    b: np.ndarray | None = None
    walk_steps: np.ndarray | None = None

    @property
    def limit_hits(self) -> np.ndarray:
        return self.caps


def _dtype_of(ct):
    if ct is C.c_double:
        return np.float64
    if ct is C.c_int64:
        return np.int64
    if ct is C.c_int32:
        return np.int32
    return (_dtype_of(ct._type_), (ct._length_,))  # fixed-size array field


def _result_dtype() -> np.dtype:
    fields = []
    for name, ct in _capi.SiteResult._fields_:
        d = _dtype_of(ct)
        fields.append((name, *d) if isinstance(d, tuple) else (name, d))
    dt = np.dtype(fields)
    assert dt.itemsize == C.sizeof(_capi.SiteResult)
    return dt


# SiteResult as a numpy record (the layout twos.h declares).
RESULT_DTYPE = _result_dtype()


def _result_buffer(n: int) -> np.ndarray:
    """Zeroed output records for n sites."""
    return np.zeros(max(n, 1), dtype=RESULT_DTYPE)


def _result(out, n, stats: dict, gradient: bool = False) -> Result:
    """Result from n SiteResult records: a RESULT_DTYPE array or any buffer
    holding them; each field is copied out contiguous."""
    rec = (out if isinstance(out, np.ndarray) else np.frombuffer(out, dtype=RESULT_DTYPE, count=n))[:n]
    f = {k: np.ascontiguousarray(rec[k]) for k in
         ("u", "v", "score", "score_se", "td", "same", "walks", "steps", "caps", "status")}
    res = Result(**f, stats=stats)
    if gradient:
        res.grad_u, res.grad_v = np.ascontiguousarray(rec["grad_u"]), np.ascontiguousarray(rec["grad_v"])
        res.grad_dot, res.grad_dot_se = np.ascontiguousarray(rec["grad_dot"]), np.ascontiguousarray(rec["grad_dot_se"])
    return res


RESULT_BYTES = C.sizeof(_capi.SiteResult)


def results_from_bytes(buf, n: int) -> Result:
    """Result of n SiteResult records read back from device memory (bytes)."""
    raw = (_capi.SiteResult * max(n, 1)).from_buffer_copy(bytes(buf)[: RESULT_BYTES * n].ljust(RESULT_BYTES, b"\0"))
    return _result(raw, n, {})


def _stats(st: _capi.Stats) -> dict:
    return {f[0]: getattr(st, f[0]) for f in _capi.Stats._fields_}


def _params(combine=SQUARE, rng=RNG_PHILOX, precision=32, schedule=PERSISTENT, seed=0, repeat=0,
            max_steps=512, eps=5e-4, walks=16, walks_max=0, ci_rel=0.02, ci_z=1.96, use_graph=False,
            block=0, waves_per_simd=0, site_offset=0, flags=0, grid_kappa=0.0, lds_stack=0,
            step_batch=0, refill_batch=0) -> _capi.Params:
    return _capi.Params(combine=combine, rng=rng, precision=precision, schedule=schedule, seed=seed,
                        repeat=repeat, max_steps=max_steps, eps=eps, walks=walks, walks_max=walks_max,
                        ci_rel=ci_rel, ci_z=ci_z, use_graph=int(use_graph), block=block,
                        waves_per_simd=waves_per_simd, site_offset=site_offset, flags=flags,
                        grid_kappa=grid_kappa, lds_stack=lds_stack, step_batch=step_batch,
                        refill_batch=refill_batch)


class Context:
    """One MI355X (HIP device), its stream, resident meshes/geometries and the
    round engine."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        _check(_lib().twos_create(device, hbm_budget, C.byref(h)))
        self._h = h
        self.device = device
        self._geom_cavities: dict[int, tuple] = {}  # geometry id -> (cavity count, cavity radius)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().twos_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- geometry -----------------------------------------------------------
    def mesh(self, vertices, faces) -> Mesh:
        """Upload a triangle mesh and build its LBVH on the GPU (the
        EmbreeMesh constructor's role, wih_twos.cpp:298-325)."""
        v = _f64(vertices).reshape(-1, 3)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        mid = C.c_int32()
        ms = C.c_double()
        _check(_lib().twos_mesh_create(self._h, _ptr(v, _capi._dp), len(v), _ptr(f, _capi._u32p), len(f),
                                       C.byref(mid), C.byref(ms)))
        depth, nodes, nbytes = C.c_int32(), C.c_int64(), C.c_int64()
        _check(_lib().twos_mesh_info(self._h, mid.value, C.byref(depth), C.byref(nodes), C.byref(nbytes)))
        return Mesh(mid.value, len(v), len(f), ms.value, depth.value, nbytes.value)

    def set_distance_grid(self, mesh: Mesh | int, cells: int, corners: int = 0) -> float:
        """Rebuild the mesh's distance grid with `cells` cells along its longest
        axis (8..1024), sampled at cell centres (corners=0), at cell corners
        (1, eight-corner bound) or as per-cell binary16 corner records (2);
        returns the build time in ms."""
        ms = C.c_double()
        _check(_lib().twos_mesh_set_distance_grid(self._h, mesh.id if isinstance(mesh, Mesh) else int(mesh),
                                                  int(cells), int(corners), C.byref(ms)))
        return ms.value

    def geometry(self, container: int, *, mesh: Mesh | int | None = None, ball=(0.0, 0.0, 0.0, 1.0),
                 box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0), spheres=None, cylinders=None, cavities=None,
                 cavity_radius: float = 0.0, mesh_radius_f32: bool = False) -> int:
        """Domain = container (ball, box or mesh interior) minus sphere and
        axis-aligned cylinder holes minus cavities of one radius (the
        reference's --cavities/--cavity-radius, wih_twos.cpp:585-626)."""
        d = _capi.GeometryDesc()
        d.container = container
        d.mesh = (mesh.id if isinstance(mesh, Mesh) else int(mesh)) if mesh is not None else -1
        d.mesh_radius_f32 = int(mesh_radius_f32)
        d.ball = (C.c_double * 4)(*ball)
        d.box_lo = (C.c_double * 3)(*box_lo)
        d.box_hi = (C.c_double * 3)(*box_hi)
        d.cavity_radius = cavity_radius
        keep = []
        for name, cnt, width in (("spheres", "n_spheres", 4), ("cylinders", "n_cylinders", 4),
                                 ("cavities", "n_cavities", 3)):
            arr = {"spheres": spheres, "cylinders": cylinders, "cavities": cavities}[name]
            if arr is not None and len(arr):
                a = _f64(arr).reshape(-1, width)
                keep.append(a)
                setattr(d, name, _ptr(a, _capi._dp))
                setattr(d, cnt, len(a))
        gid = C.c_int32()
        _check(_lib().twos_geometry_create(self._h, C.byref(d), C.byref(gid)))
        self._geom_cavities[gid.value] = (int(d.n_cavities), float(cavity_radius))
        return gid.value

    def union_sphere(self, geom: int, center, radius: float) -> int:
        """Derived geometry with one more spherical cavity (fresh-seed
        insertion validation)."""
        c = _f64(center)
        gid = C.c_int32()
        _check(_lib().twos_geometry_union_sphere(self._h, geom, _ptr(c, _capi._dp), radius, C.byref(gid)))
        # A sphere of the base's cavity radius (or the first one) joins its
        # cavities; another radius becomes a sphere hole.
        k, r = self._geom_cavities.get(geom, (0, 0.0))
        self._geom_cavities[gid.value] = (k + 1, radius) if k == 0 or r == radius else (k, r)
        return gid.value

    def release_geometry(self, geom: int) -> None:
        _check(_lib().twos_geometry_release(self._h, geom))
        self._geom_cavities.pop(geom, None)

    def memory(self) -> dict:
        """Device bytes the context accounts for (resident geometry, engines),
        its HBM budget and the bytes one admitted site costs; a server admits
        a request only while geometry + engines + its sites fit the budget."""
        g, e, b, s = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(_lib().twos_memory_info(self._h, C.byref(g), C.byref(e), C.byref(b), C.byref(s)))
        return {"geometry_bytes": g.value, "engine_bytes": e.value, "budget_bytes": b.value, "slot_bytes": s.value}

    def release_mesh(self, mesh: Mesh | int) -> None:
        _check(_lib().twos_mesh_release(self._h, mesh.id if isinstance(mesh, Mesh) else int(mesh)))

    # -- queries ------------------------------------------------------------
    def inside(self, mesh: Mesh | int, points) -> np.ndarray:
        """Ray-parity membership (EmbreeMesh::inside, wih_twos.cpp:361-394)."""
        p = _f64(points).reshape(-1, 3)
        out = np.zeros(len(p), dtype=np.int32)
        mid = mesh.id if isinstance(mesh, Mesh) else int(mesh)
        _check(_lib().twos_inside(self._h, mid, _ptr(p, _capi._dp), len(p), _ptr(out, _capi._i32p)))
        return out.astype(bool)

    def distance_bound(self, geom: int, points, precision: int = 32) -> np.ndarray:
        """The distance grid's lower bound of the mesh distance at each point
        (-inf without a grid)."""
        p = _f64(points).reshape(-1, 3)
        out = np.zeros(len(p))
        _check(_lib().twos_distance_bound(self._h, geom, precision, _ptr(p, _capi._dp), len(p),
                                          _ptr(out, _capi._dp)))
        return out

    def closest(self, geom: int, points, precision: int = 64):
        """Distance to the domain boundary, closest point and boundary class."""
        p = _f64(points).reshape(-1, 3)
        n = len(p)
        d = np.zeros(n)
        q = np.zeros((n, 3))
        c = np.zeros(n, dtype=np.int32)
        _check(_lib().twos_closest(self._h, geom, precision, _ptr(p, _capi._dp), n, _ptr(d, _capi._dp),
                                   _ptr(q, _capi._dp), _ptr(c, _capi._i32p)))
        return d, q, c

    def trace(self, geom: int, start, side_: Side, *, rng=RNG_MT_KEY, seed=0, site=0, walk=0, repeat=0,
              eps=5e-4, max_steps=512, precision=64) -> np.ndarray:
        """Rows (x, y, z, radius) of one path (write_trace,
        wih_twos.cpp:1725-1744).  With RNG_MT_KEY the stream is
        mt19937_64(splitmix64(seed + site))."""
        s = _f64(start)
        rows = np.zeros((max_steps + 2, 4))
        n = C.c_int32()
        prm = _params(rng=rng, precision=precision, seed=seed, repeat=repeat, eps=eps, max_steps=max_steps,
                      walks=2, schedule=PLAIN)
        _check(_lib().twos_trace(self._h, geom, C.byref(side_), C.byref(prm), _ptr(s, _capi._dp), site, walk,
                                 _ptr(rows, _capi._dp), len(rows), C.byref(n)))
        return rows[:n.value].copy()

    # Per-site counters of differential(), in column order.
    DIFF_COUNTS = ("outer_steps", "nested_steps", "cavity_terminations", "near_seam", "segment_unsafe", "unsafe",
                   "limits", "distance_queries", "parity_rays")

    def differential(self, geom: int, sites, side_: Side, *, seed=0, repeat=0, walks=16, eps=1e-4,
                     max_steps=2048, offset=0.004, site_offset=0, rng=RNG_MT_KEY, precision=64,
                     distance_grid=False, grid_kappa=0.0):
        """Differential WoS cavity derivatives (evaluate_differential_site,
        wih_twos.cpp:725-928): (values [n][n_cavities][8] = centre derivative
        x y z, their SEs, radius derivative, its SE; counts [n][9] in
        DIFF_COUNTS order).  rng RNG_MT_KEY replays the reference's
        mt19937_64 streams in binary64; RNG_PHILOX is the production form
        (one persistent launch, precision 32 or 64, optionally the mesh's
        distance grid for steps at least grid_kappa cells from the mesh)."""
        p = _f64(sites).reshape(-1, 3)
        n = len(p)
        k = self._geom_cavities.get(geom, (0, 0.0))[0]
        out = np.zeros((n, max(k, 1), 8))
        counts = np.zeros((n, 9), dtype=np.int64)
        prm = _params(rng=rng, precision=precision, seed=seed, repeat=repeat, eps=eps, max_steps=max_steps,
                      walks=walks, schedule=PLAIN, site_offset=site_offset,
                      flags=FLAG_DISTANCE_GRID if distance_grid else 0, grid_kappa=grid_kappa)
        _check(_lib().twos_differential(self._h, geom, C.byref(side_), C.byref(prm), _ptr(p, _capi._dp), n,
                                        float(offset), _ptr(out, _capi._dp),
                                        counts.ctypes.data_as(C.POINTER(C.c_int64))))
        return out[:, :k], counts

    def certify(self, geom: int, sites, *, seed=0, walks=1024, max_steps=128, eps=1e-6, radius_cap=0.1,
                value_cap=0.1, source=_capi.CERT_AFFINE, slope=0.65, site_offset=0, sample_begin=0, philox=False,
                grid_cells=0, grid_kappa=0.0):
        """Certified minimum-benefit samples (cert_walk, minimum_benefit.hpp:
        122-143): (value, steps, flags), [n][walks] each, for samples
        sample_begin .. sample_begin + walks - 1; flags CERT_CAPPED,
        CERT_CLIPPED, CERT_CONTAINMENT.  On the reference's mt19937_64
        streams, or with philox=True on Philox streams in one persistent
        launch (the production form), optionally with the mesh's certified
        grid (grid_cells along its longest axis) for steps at least
        grid_kappa cells (0: 2) from the mesh."""
        p = _f64(sites).reshape(-1, 3)
        n = len(p)
        prm = _capi.CertParams(seed=seed, walks=walks, max_steps=max_steps, eps=eps, radius_cap=radius_cap,
                               value_cap=value_cap, slope=slope, source=source, site_offset=site_offset,
                               sample_begin=sample_begin, philox=int(bool(philox)), grid_cells=grid_cells,
                               grid_kappa=grid_kappa)
        value = np.zeros((n, walks))
        steps = np.zeros((n, walks), dtype=np.int32)
        flags = np.zeros((n, walks), dtype=np.int32)
        i32 = C.POINTER(C.c_int32)
        _check(_lib().twos_certify(self._h, geom, C.byref(prm), _ptr(p, _capi._dp), n, _ptr(value, _capi._dp),
                                   steps.ctypes.data_as(i32), flags.ctypes.data_as(i32)))
        return value, steps, flags

    def cert_distance(self, geom: int, points) -> np.ndarray:
        """Certified lower bound of the domain distance (cert_domain_distance)."""
        p = _f64(points).reshape(-1, 3)
        out = np.zeros(len(p))
        _check(_lib().twos_cert_distance(self._h, geom, _ptr(p, _capi._dp), len(p), _ptr(out, _capi._dp)))
        return out

    def cert_grid_lower(self, mesh, cells: int, points) -> np.ndarray:
        """The certified grid's lower bound of the mesh's box-union distance
        (built with `cells` cells along the mesh's longest axis if needed)."""
        p = _f64(points).reshape(-1, 3)
        out = np.zeros(len(p))
        _check(_lib().twos_cert_grid_lower(self._h, mesh.id if isinstance(mesh, Mesh) else int(mesh), cells,
                                           _ptr(p, _capi._dp), len(p), _ptr(out, _capi._dp)))
        return out

    # -- the estimator ------------------------------------------------------
    def estimate(self, geom: int, sites, sides, *, combine=SQUARE, rng=RNG_PHILOX, precision=32,
                 schedule=PERSISTENT, seed=0, repeat=0, walks=16, walks_max=0, ci_rel=0.02, ci_z=1.96,
                 eps=5e-4, max_steps=512, use_graph=False, block=0, waves_per_simd=0, site_offset=0,
                 count_traversal=False, distance_grid=False, grid_kappa=0.0, lds_stack=0, gradient=False,
                 step_batch=0, refill_batch=0, dump: bool = False, raise_on_cap: bool = True) -> Result:
        """Per-site TWoS estimate.  walks_max > walks doubles each site's walks
        until ci_z * score_se <= ci_rel * |score| or walks_max is reached.
        distance_grid: mesh steps far from the boundary take a conservative
        radius from the mesh's distance grid (unbiased; paths differ from the
        exact walk) when it is at least grid_kappa cell sizes (0: 4).
        gradient: also estimate grad u, grad v and their cross-index product
        from each path's first step (Result.grad_*)."""
        p = _f64(sites).reshape(-1, 3)
        n = len(p)
        sd = (Side * 2)(sides[0], sides[1] if len(sides) > 1 else sides[0])
        prm = _params(combine=combine, rng=rng, precision=precision, schedule=schedule, seed=seed,
                      repeat=repeat, max_steps=max_steps, eps=eps, walks=walks, walks_max=walks_max,
                      ci_rel=ci_rel, ci_z=ci_z, use_graph=use_graph, block=block,
                      waves_per_simd=waves_per_simd, site_offset=site_offset,
                      flags=(FLAG_COUNT_TRAVERSAL if count_traversal else 0)
                      | (FLAG_DISTANCE_GRID if distance_grid else 0) | (FLAG_GRADIENT if gradient else 0),
                      grid_kappa=grid_kappa, lds_stack=lds_stack, step_batch=step_batch,
                      refill_batch=refill_batch)
        out = _result_buffer(n)
        out_p = out.ctypes.data_as(C.POINTER(_capi.SiteResult))
        st = _capi.Stats()
        if dump:
            stride = max(walks, walks_max)
            da = np.zeros((n, stride))
            db = np.zeros((n, stride))
            ds = np.zeros((n, stride), dtype=np.int32)
            status = _lib().twos_estimate_dump(self._h, geom, sd, C.byref(prm), _ptr(p, _capi._dp), n, out_p,
                                               C.byref(st), _ptr(da, _capi._dp), _ptr(db, _capi._dp),
                                               _ptr(ds, _capi._i32p), stride)
        else:
            status = _lib().twos_estimate(self._h, geom, sd, C.byref(prm), _ptr(p, _capi._dp), n, out_p,
                                          C.byref(st))
        res = _result(out, n, _stats(st), gradient)
        if dump:
            res.a, res.b, res.walk_steps = da, db, ds
        if status == _capi.ERR_STEP_CAP and not raise_on_cap:
            return res
        _check(status, res)
        return res

    def estimate_device(self, geom: int, d_sites: int, n: int, d_out: int, sides, *, stream: int = 0,
                        static: bool = False, combine=SQUARE, rng=RNG_PHILOX, precision=32, schedule=PERSISTENT,
                        seed=0, repeat=0, walks=16, walks_max=0, ci_rel=0.02, ci_z=1.96, eps=5e-4, max_steps=512,
                        block=0, waves_per_simd=0, site_offset=0, count_traversal=False, distance_grid=False,
                        grid_kappa=0.0, lds_stack=0, step_batch=0, refill_batch=0, use_graph=False,
                        gradient=False, raise_on_cap: bool = True) -> dict:
        """estimate() on device memory: d_sites is the device address of n x 3
        binary64 positions, d_out of n SiteResult records (RESULT_BYTES each),
        both ordered on `stream` (a hipStream_t handle, 0 for the null
        stream): the estimate starts after the work queued there and later
        work there sees the results.  static=True runs the allocation-free
        variant (every buffer in the device's static scratch array; no
        graphs or gradients, walks per side up to its per-round limit).
        Returns the stats dict; decode d_out with results_from_bytes."""
        sd = (Side * 2)(sides[0], sides[1] if len(sides) > 1 else sides[0])
        prm = _params(combine=combine, rng=rng, precision=precision, schedule=schedule, seed=seed, repeat=repeat,
                      max_steps=max_steps, eps=eps, walks=walks, walks_max=walks_max, ci_rel=ci_rel, ci_z=ci_z,
                      block=block, waves_per_simd=waves_per_simd, site_offset=site_offset,
                      flags=(FLAG_COUNT_TRAVERSAL if count_traversal else 0)
                      | (FLAG_DISTANCE_GRID if distance_grid else 0) | (FLAG_GRADIENT if gradient else 0),
                      grid_kappa=grid_kappa, lds_stack=lds_stack, step_batch=step_batch, refill_batch=refill_batch,
                      use_graph=use_graph)
        st = _capi.Stats()
        fn = _lib().twos_estimate_device_static if static else _lib().twos_estimate_device
        status = fn(self._h, geom, sd, C.byref(prm), C.c_void_p(d_sites), n, C.c_void_p(d_out), C.byref(st),
                    C.c_void_p(stream))
        stats = _stats(st)
        if not (status == _capi.ERR_STEP_CAP and not raise_on_cap):
            _check(status)
        return stats

    def server(self, max_sites: int, max_walks_per_round: int) -> "Server":
        return Server(self, max_sites, max_walks_per_round)


class Server:
    """Continuous-batching request loop over one context (twos_server_*).
    Requests are admitted between rounds as slots free up; all admitted sites
    share each persistent launch."""

    def __init__(self, ctx: Context, max_sites: int, max_walks_per_round: int):
        h = C.c_void_p()
        _check(_lib().twos_server_create(ctx._h, max_sites, max_walks_per_round, C.byref(h)))
        self._h = h
        self.ctx = ctx
        self._pending: dict[int, tuple] = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib().twos_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, geom: int, sites, sides, *, combine=CROSS, precision=32, seed=0, walks=256,
               walks_max=0, ci_rel=0.02, ci_z=1.96, eps=1e-4, max_steps=4096, use_graph=False,
               waves_per_simd=0, site_offset=0, distance_grid=False, grid_kappa=0.0, block=0,
               lds_stack=0, step_batch=0, refill_batch=0) -> int:
        p = _f64(sites).reshape(-1, 3)
        n = len(p)
        sd = (Side * 2)(sides[0], sides[1] if len(sides) > 1 else sides[0])
        prm = _params(combine=combine, precision=precision, schedule=PERSISTENT, seed=seed, walks=walks,
                      walks_max=walks_max, ci_rel=ci_rel, ci_z=ci_z, eps=eps, max_steps=max_steps,
                      use_graph=use_graph, waves_per_simd=waves_per_simd, site_offset=site_offset,
                      flags=FLAG_DISTANCE_GRID if distance_grid else 0, grid_kappa=grid_kappa, block=block,
                      lds_stack=lds_stack, step_batch=step_batch, refill_batch=refill_batch)
        out = (_capi.SiteResult * n)()
        rid = C.c_int64()
        _check(_lib().twos_server_submit(self._h, geom, sd, C.byref(prm), _ptr(p, _capi._dp), n, out,
                                         C.byref(rid)))
        self._pending[rid.value] = (out, n, p, sd)
        return rid.value

    def step(self) -> dict[int, Result]:
        """Run one round; returns the requests that ended in it.  Each
        result's stats["status"] is 0, ERR_STEP_CAP (a path hit max_steps;
        results written) or the error that kept the request from running
        (ERR_ADMISSION); the other requests go on either way."""
        cap = max(1, len(self._pending))
        ids = (C.c_int64 * cap)()
        codes = (C.c_int32 * cap)()
        nd = C.c_int32()
        _check(_lib().twos_server_step_status(self._h, ids, codes, cap, C.byref(nd)))
        done = {}
        for k in range(min(nd.value, cap)):
            out, n, _, _ = self._pending.pop(ids[k])
            done[ids[k]] = _result(out, n, {"status": int(codes[k])})
        return done

    def pending(self) -> tuple[int, int]:
        q, a = C.c_int64(), C.c_int64()
        _check(_lib().twos_server_pending(self._h, C.byref(q), C.byref(a)))
        return q.value, a.value
