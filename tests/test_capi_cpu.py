"""The C-ABI library on the CPU: it loads, exports every entry point that
include/twos.h declares, and its Python binding covers exactly that set.
No compute is called (there is no GPU here)."""

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "twos.h"


def declared() -> set[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return set(re.findall(r"\b(twos_[a-z_]+)\s*\(", text))


def test_header_declares_the_boundary():
    names = declared()
    for required in ("twos_create", "twos_destroy", "twos_estimate", "twos_mesh_create",
                     "twos_geometry_create", "twos_inside", "twos_last_error", "twos_server_submit"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from twos import _capi
    lib = _capi.load()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_capi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (twos_[a-z_]+)$", out, flags=re.M))
    assert declared() <= exported


def test_binding_matches_header():
    from twos import _capi
    assert set(_capi.SIGNATURES) == declared()


def test_struct_layouts_match_header_sizes(tmp_path):
    """The ctypes mirrors have the size and field offsets the C compiler gives
    the header's structs."""
    import subprocess
    from twos import _capi
    structs = {"twos_side": _capi.Side, "twos_site_result": _capi.SiteResult, "twos_params": _capi.Params,
               "twos_geometry_desc": _capi.GeometryDesc, "twos_stats": _capi.Stats,
               "twos_cert_params": _capi.CertParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT / "include" / "twos.h"}"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, *_ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, what, value = line.split()
        py = structs[cname]
        want = ctypes.sizeof(py) if what == "size" else getattr(py, what).offset
        assert int(value) == want, (cname, what)


def test_version_and_no_device_error_without_gpu():
    import twos
    assert "gfx950" in twos.version()
    try:
        ctx = twos.Context()
    except twos.DeviceError as e:
        assert "device" in str(e).lower()
    else:  # on a GPU box the context simply opens
        ctx.close()


def test_result_records_decode_to_contiguous_fields():
    """Estimate results arrive as SiteResult records (twos.h): a record array
    and the same records as raw bytes (the device entry point's read-back)
    decode to the same contiguous per-site fields, gradient fields included,
    and zero sites to empty arrays."""
    import twos
    n = 37
    rec = np.zeros(n, dtype=twos.RESULT_DTYPE)
    assert twos.RESULT_DTYPE.itemsize == twos.RESULT_BYTES
    rng = np.random.default_rng(4)
    for k in ("u", "v", "score", "score_se", "td", "same", "grad_dot", "grad_dot_se"):
        rec[k] = rng.normal(size=n)
    rec["grad_u"] = rng.normal(size=(n, 3))
    rec["walks"] = rng.integers(2, 1 << 40, n)
    rec["steps"] = rng.integers(0, 1 << 50, n)
    rec["caps"] = rng.integers(0, 9, n)
    rec["status"] = rng.integers(0, 2, n)
    a = twos._result(rec, n, {"x": 1}, gradient=True)
    b = twos.results_from_bytes(rec.tobytes(), n)
    for k in ("u", "v", "score", "score_se", "td", "same", "walks", "steps", "caps", "status"):
        assert getattr(a, k).flags["C_CONTIGUOUS"], k
        np.testing.assert_array_equal(getattr(a, k), rec[k], err_msg=k)
        np.testing.assert_array_equal(getattr(b, k), rec[k], err_msg=k)
    np.testing.assert_array_equal(a.grad_u, rec["grad_u"])
    np.testing.assert_array_equal(a.grad_dot_se, rec["grad_dot_se"])
    assert a.stats == {"x": 1} and b.grad_u is None
    rec["u"][0] = 99.0  # the fields are copies, not views of the records
    assert rec["u"][0] == 99.0 and a.u[0] != 99.0
    empty = twos._result(twos._result_buffer(0), 0, {})
    assert len(empty.u) == 0 and len(empty.walks) == 0


def test_side_validation():
    import twos
    with pytest.raises(twos.InvalidArgument):
        twos.side(poly=[1, 2, 3])
    with pytest.raises(twos.InvalidArgument):
        twos.side(poly_classes=(9,))
    s = twos.reference_affine_side(0.65)
    assert s.source_kind == 1 and list(s.src) == [1.0, 0.0, 0.65, 0.0, 0.0, 0.5, 0.0]


def test_workloads_are_deterministic_and_sized():
    from twos import workloads as W
    a, b = W.sh_icosphere(n_sites=512), W.sh_icosphere(n_sites=512)
    np.testing.assert_array_equal(a.sites, b.sites)
    assert a.faces.shape == (81920, 3) and a.vertices.shape == (40962, 3)
    assert a.vertices.min() >= 0.0 and a.vertices.max() <= 1.0 + 1e-12
    c = W.cooling_block()
    assert 2500 < len(c.sites) < 4096
    assert len(c.spheres) == 12 and len(c.cylinders) == 2
    u = W.unit_ball()
    assert u.sites.shape == (64, 3) and np.linalg.norm(u.sites, axis=1).max() <= 0.8
