"""ctypes binding of include/twos.h (libtwos_hip.so, built for gfx950).

The library is the product path: there is no CPU fallback.  Loading fails
loudly when the shared object is missing or was built for another layout.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("TWOS_LIBRARY", HERE / "libtwos_hip.so"))

OK = 0
ERR_INVALID_ARGUMENT = 2
ERR_STEP_CAP = 3
ERR_GEOMETRY = 4
ERR_ADMISSION = 5
ERR_DEVICE = 6
ERR_NOT_FOUND = 7

BALL, BOX, MESH = 0, 1, 2
SQUARE, PAIRED, CROSS, COUPLED = 0, 1, 2, 3
RNG_PHILOX, RNG_MT_SITE, RNG_MT_PAIRED, RNG_MT_KEY, RNG_MT_FD = 0, 1, 2, 3, 4
PLAIN, PERSISTENT = 0, 1
FLAG_COUNT_TRAVERSAL = 1
FLAG_DISTANCE_GRID = 2
FLAG_GRADIENT = 4
TERM_BUMP, TERM_UNIT, TERM_GAUSS = 0, 1, 2
MAX_TERMS = 8

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)


class Side(C.Structure):
    _fields_ = [("poly", C.c_double * 10), ("cls_const", C.c_double * 8), ("cls_poly_mask", C.c_uint32),
                ("source_kind", C.c_int32), ("src", C.c_double * 7), ("n_terms", C.c_int32), ("pad", C.c_int32),
                ("terms", (C.c_double * 8) * MAX_TERMS)]


class GeometryDesc(C.Structure):
    _fields_ = [("container", C.c_int32), ("n_spheres", C.c_int32), ("n_cylinders", C.c_int32),
                ("n_cavities", C.c_int32), ("mesh", C.c_int32), ("mesh_radius_f32", C.c_int32),
                ("ball", C.c_double * 4), ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3),
                ("cavity_radius", C.c_double), ("spheres", _dp), ("cylinders", _dp), ("cavities", _dp)]


class CertParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("walks", C.c_int32), ("max_steps", C.c_int32), ("eps", C.c_double),
                ("radius_cap", C.c_double), ("value_cap", C.c_double), ("slope", C.c_double),
                ("source", C.c_int32), ("philox", C.c_int32), ("site_offset", C.c_int64),
                ("sample_begin", C.c_int64), ("grid_cells", C.c_int32), ("grid_kappa", C.c_float)]


CERT_AFFINE, CERT_HOUSING = 0, 1
CERT_CAPPED, CERT_CLIPPED, CERT_CONTAINMENT = 1, 2, 4


class Params(C.Structure):
    _fields_ = [("combine", C.c_int32), ("rng", C.c_int32), ("precision", C.c_int32),
                ("schedule", C.c_int32), ("seed", C.c_uint64), ("repeat", C.c_int32),
                ("max_steps", C.c_int32), ("eps", C.c_double), ("walks", C.c_int32),
                ("walks_max", C.c_int32), ("ci_rel", C.c_double), ("ci_z", C.c_double),
                ("use_graph", C.c_int32), ("block", C.c_int32), ("waves_per_simd", C.c_int32),
                ("site_offset", C.c_int32), ("flags", C.c_int32), ("grid_kappa", C.c_float),
                ("lds_stack", C.c_int32), ("step_batch", C.c_int32), ("refill_batch", C.c_int32)]


class SiteResult(C.Structure):
    _fields_ = [("u", C.c_double), ("v", C.c_double), ("score", C.c_double), ("score_se", C.c_double),
                ("td", C.c_double), ("same", C.c_double), ("walks", C.c_int64), ("steps", C.c_int64),
                ("caps", C.c_int64), ("status", C.c_int32), ("pad", C.c_int32),
                ("grad_u", C.c_double * 3), ("grad_v", C.c_double * 3), ("grad_dot", C.c_double),
                ("grad_dot_se", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("walk_kernel_ms", C.c_double), ("reduce_kernel_ms", C.c_double),
                ("walk_launches", C.c_int64), ("steps", C.c_int64), ("paths", C.c_int64),
                ("nodes_visited", C.c_int64), ("tris_tested", C.c_int64), ("rounds", C.c_int32),
                ("block", C.c_int32), ("waves_per_simd", C.c_int32), ("grid", C.c_int32),
                ("grid_lookups", C.c_int64), ("grid_hits", C.c_int64),
                ("top_nodes_visited", C.c_int64)]


_ctx_p = C.c_void_p
_srv_p = C.c_void_p

# name -> (restype, argtypes); the set of names is exactly the extern "C"
# entry points declared in include/twos.h.
SIGNATURES = {
    "twos_version": (C.c_char_p, []),
    "twos_last_error": (C.c_char_p, []),
    "twos_create": (C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(_ctx_p)]),
    "twos_destroy": (None, [_ctx_p]),
    "twos_mesh_create": (C.c_int32, [_ctx_p, _dp, C.c_int64, _u32p, C.c_int64, _i32p, _dp]),
    "twos_mesh_info": (C.c_int32, [_ctx_p, C.c_int32, _i32p, _i64p, _i64p]),
    "twos_mesh_set_distance_grid": (C.c_int32, [_ctx_p, C.c_int32, C.c_int32, C.c_int32, _dp]),
    "twos_mesh_release": (C.c_int32, [_ctx_p, C.c_int32]),
    "twos_memory_info": (C.c_int32, [_ctx_p, _i64p, _i64p, _i64p, _i64p]),
    "twos_geometry_create": (C.c_int32, [_ctx_p, C.POINTER(GeometryDesc), _i32p]),
    "twos_geometry_union_sphere": (C.c_int32, [_ctx_p, C.c_int32, _dp, C.c_double, _i32p]),
    "twos_geometry_release": (C.c_int32, [_ctx_p, C.c_int32]),
    "twos_estimate": (C.c_int32, [_ctx_p, C.c_int32, C.POINTER(Side), C.POINTER(Params), _dp, C.c_int64,
                                  C.POINTER(SiteResult), C.POINTER(Stats)]),
    "twos_estimate_device": (C.c_int32, [_ctx_p, C.c_int32, C.POINTER(Side), C.POINTER(Params), C.c_void_p,
                                         C.c_int64, C.c_void_p, C.POINTER(Stats), C.c_void_p]),
    "twos_estimate_device_static": (C.c_int32, [_ctx_p, C.c_int32, C.POINTER(Side), C.POINTER(Params), C.c_void_p,
                                                C.c_int64, C.c_void_p, C.POINTER(Stats), C.c_void_p]),
    "twos_estimate_dump": (C.c_int32, [_ctx_p, C.c_int32, C.POINTER(Side), C.POINTER(Params), _dp, C.c_int64,
                                       C.POINTER(SiteResult), C.POINTER(Stats), _dp, _dp, _i32p, C.c_int64]),
    "twos_inside": (C.c_int32, [_ctx_p, C.c_int32, _dp, C.c_int64, _i32p]),
    "twos_closest": (C.c_int32, [_ctx_p, C.c_int32, C.c_int32, _dp, C.c_int64, _dp, _dp, _i32p]),
    "twos_distance_bound": (C.c_int32, [_ctx_p, C.c_int32, C.c_int32, _dp, C.c_int64, _dp]),
    "twos_differential": (C.c_int32, [_ctx_p, C.c_int32, C.POINTER(Side), C.POINTER(Params), _dp, C.c_int64,
                                      C.c_double, _dp, C.POINTER(C.c_int64)]),
    "twos_certify": (C.c_int32, [_ctx_p, C.c_int32, C.POINTER(CertParams), _dp, C.c_int64, _dp, _i32p, _i32p]),
    "twos_cert_distance": (C.c_int32, [_ctx_p, C.c_int32, _dp, C.c_int64, _dp]),
    "twos_cert_grid_lower": (C.c_int32, [_ctx_p, C.c_int32, C.c_int32, _dp, C.c_int64, _dp]),
    "twos_trace": (C.c_int32, [_ctx_p, C.c_int32, C.POINTER(Side), C.POINTER(Params), _dp, C.c_int64,
                               C.c_int32, _dp, C.c_int32, _i32p]),
    "twos_server_create": (C.c_int32, [_ctx_p, C.c_int64, C.c_int64, C.POINTER(_srv_p)]),
    "twos_server_destroy": (None, [_srv_p]),
    "twos_server_submit": (C.c_int32, [_srv_p, C.c_int32, C.POINTER(Side), C.POINTER(Params), _dp, C.c_int64,
                                       C.POINTER(SiteResult), _i64p]),
    "twos_server_step": (C.c_int32, [_srv_p, _i64p, C.c_int32, _i32p]),
    "twos_server_step_status": (C.c_int32, [_srv_p, _i64p, _i32p, C.c_int32, _i32p]),
    "twos_server_pending": (C.c_int32, [_srv_p, _i64p, _i64p]),
}

_lib = None


def load() -> C.CDLL:
    """Load libtwos_hip.so; raises if it is absent (build it with
    `python __graft_entry__.py` or `make -C walking-into-holes_amd`)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: the HIP library must be built for gfx950 "
                              "(make -C walking-into-holes_amd); there is no CPU fallback")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
