"""Certified minimum-benefit sampling (twos_certify; cert_walk and
run_minimum_benefit, benchmarks/compiled_bunny/src/minimum_benefit.hpp:
122-188) on an MI355X, through the C-ABI, against the pilot's retained
--mode validate samples (benchmarks/minimum_benefit/pilot-floor-policy/).

- every retained attempt (six on the watertight bunny with the affine
  source, six on the housing with its chips; up to five cavities): the first
  2,048 samples' step counts and capped / clipped flags exactly and values
  to 1e-10 (bitwise on most: the device library's cos / sin may differ from
  glibc's by an ulp, and over up to 128 steps that moves a few values by
  about 1e-12 relative);
- all 65,536 samples of every attempt: each stage's path_caps and clipped
  counts exactly, its mean to 1e-11 (summed in long double, as the
  reference does);
- the certified clearance of every finalist bitwise;
- a later block of samples (sample_begin) equals the same samples of one
  large call, so stages can be run one at a time;
- the production form (Philox streams, one persistent launch): every
  sample's steps and flags equal to the oracle's Philox form, values to
  1e-10; results keyed by (seed, site, sample);
- the certified grid: its bound never above the exact certified distance,
  and walks that use it agree with the exact walk within SE.
"""

import numpy as np
import pytest

import cert_fixtures as CF

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def setup(ctx, golden):
    meshes = {k: ctx.mesh(*CF.mesh(golden, k)) for k in ("bunny", "housing")}
    return meshes, CF.load(golden)


def _certify(ctx, meshes, a, walks, sample_begin=0, points=None, site_offset=0, philox=False, grid_cells=0):
    g = ctx.geometry(twos.MESH, mesh=meshes[a.kind], cavities=a.holes if len(a.holes) else None,
                     cavity_radius=a.radius)
    try:
        return ctx.certify(g, a.point if points is None else points, seed=a.seed, walks=walks,
                           max_steps=a.max_steps, eps=a.epsilon, radius_cap=a.radius_cap, value_cap=a.value_cap,
                           source=twos.CERT_AFFINE if a.kind == "bunny" else twos.CERT_HOUSING, slope=a.slope,
                           sample_begin=sample_begin, site_offset=site_offset, philox=philox,
                           grid_cells=grid_cells)
    finally:
        ctx.release_geometry(g)


@pytest.fixture(scope="module")
def oracle_meshes(golden):
    O = pytest.importorskip("oracle")
    return {k: O.Mesh(*CF.mesh(golden, k)) for k in ("bunny", "housing")}


@pytest.mark.parametrize("index", [0, 7])
def test_philox_samples_match_oracle(ctx, setup, oracle_meshes, index):
    """The production form (Philox streams, one persistent launch) against the
    oracle's Philox form on a retained finalist (the bunny with the affine
    source; the housing with its chips and cavities): every sample's step
    count and flags exactly, values to 1e-10 and most bitwise."""
    import oracle as O
    meshes, attempts = setup
    a = attempts[index]
    n = 4096
    print("certify philox", a.source, flush=True)
    value, steps, flags = _certify(ctx, meshes, a, n, philox=True)
    geo = O.Geometry(O.MESH, mesh=oracle_meshes[a.kind], cavities=a.holes if len(a.holes) else None,
                     cavity_radius=a.radius)
    rv, rs, rf = O.certify(geo, a.point, seed=a.seed, walks=n, max_steps=a.max_steps, eps=a.epsilon,
                           radius_cap=a.radius_cap, value_cap=a.value_cap,
                           source=O.CERT_AFFINE if a.kind == "bunny" else O.CERT_HOUSING, slope=a.slope,
                           threads=O.host_threads(), rng=O.RNG_PHILOX)
    np.testing.assert_array_equal(steps, rs)
    np.testing.assert_array_equal(flags, rf)
    np.testing.assert_allclose(value, rv, rtol=1e-10, atol=0)
    assert np.mean(value == rv) > 0.5
    assert not np.any(flags & twos.CERT_CONTAINMENT)


@pytest.mark.parametrize("kind", ["bunny", "housing"])
def test_cert_grid_is_conservative(ctx, setup, golden, kind):
    """The certified grid's bound never exceeds the exact certified box-union
    distance (twos_cert_distance on the bare mesh), in and around the mesh,
    at two resolutions; inside the grid (the mesh's bounds plus one cell) and
    far from the mesh it is within two cell diagonals of it (the grid is
    used, not just safe; outside the grid the nearest cell's corners give
    a looser bound)."""
    meshes, _ = setup
    v, _ = CF.mesh(golden, kind)
    lo, hi = v.min(axis=0), v.max(axis=0)
    rng = np.random.default_rng(11)
    pts = lo - 0.1 * (hi - lo) + rng.random((20000, 3)) * 1.2 * (hi - lo)
    g = ctx.geometry(twos.MESH, mesh=meshes[kind])
    try:
        exact = ctx.cert_distance(g, pts)
    finally:
        ctx.release_geometry(g)
    for cells in (64, 256):
        print("cert grid", kind, cells, flush=True)
        lb = ctx.cert_grid_lower(meshes[kind], cells, pts)
        assert np.all(lb <= exact), (cells, np.max(lb - exact))
        h = np.max(hi - lo) / cells
        far = (exact > 4 * h) & np.all((pts >= lo) & (pts <= hi), axis=1)
        assert far.sum() > 1000
        assert np.all(lb[far] >= exact[far] - 2 * np.sqrt(3) * h), cells


@pytest.mark.parametrize("index", [0, 7])
def test_cert_grid_walks_against_exact(ctx, setup, index):
    """Certified walks that take the grid's bound far from the mesh: other
    paths (more, smaller steps: the bound is below the exact certified
    distance), the same certified benefit and clipped rate within the
    combined SE of the exact walk, no containment failure, and no fewer walks
    reaching max_steps (smaller steps cap more walks)."""
    meshes, attempts = setup
    a = attempts[index]
    n = 65536
    print("certify philox grid", a.source, flush=True)
    ev, es, ef = _certify(ctx, meshes, a, n, philox=True)
    gv, gs, gf = _certify(ctx, meshes, a, n, philox=True, grid_cells=256)
    assert not np.array_equal(gs, es)  # the grid was used
    assert not np.any(gf & twos.CERT_CONTAINMENT)
    for x, y in ((ev[0], gv[0]), (((ef[0] >> 1) & 1).astype(float), ((gf[0] >> 1) & 1).astype(float))):
        se = np.hypot(np.std(x), np.std(y)) / np.sqrt(n)
        assert abs(np.mean(x) - np.mean(y)) <= 4 * se + 1e-15, (np.mean(x), np.mean(y), se)
    ce, cg = (ef[0] & 1).astype(float), (gf[0] & 1).astype(float)
    assert np.mean(cg) >= np.mean(ce) - 4 * np.hypot(np.std(ce), np.std(cg)) / np.sqrt(n)
    assert np.mean(gs) >= np.mean(es)


def test_philox_samples_are_keyed_by_index(ctx, setup):
    """A Philox sample depends on (seed, site, sample) only: blocks of samples
    (sample_begin) and a site offset give the records of one large call."""
    meshes, attempts = setup
    a = attempts[3]
    pts = np.vstack([a.point, a.point + [0.01, 0.0, 0.0]])
    full = _certify(ctx, meshes, a, 8192, points=pts, philox=True)
    block = _certify(ctx, meshes, a, 2048, sample_begin=4096, points=pts, philox=True)
    for x, y in zip(full, block):
        np.testing.assert_array_equal(x[:, 4096:6144], y)
    second = _certify(ctx, meshes, a, 8192, points=pts[1:], site_offset=1, philox=True)
    for x, y in zip(full, second):
        np.testing.assert_array_equal(x[1:], y)
    ref = _certify(ctx, meshes, a, 8192, points=pts)
    assert not np.array_equal(ref[0], full[0])  # the streams differ from the reference's
    gfull = _certify(ctx, meshes, a, 8192, points=pts, philox=True, grid_cells=128)
    gblock = _certify(ctx, meshes, a, 2048, sample_begin=4096, points=pts, philox=True, grid_cells=128)
    for x, y in zip(gfull, gblock):
        np.testing.assert_array_equal(x[:, 4096:6144], y)


def test_first_samples_of_every_attempt(ctx, setup):
    meshes, attempts = setup
    for a in attempts:
        print("certify", a.source, len(a.holes), "cavities", flush=True)
        value, steps, flags = _certify(ctx, meshes, a, len(a.value))
        np.testing.assert_array_equal(steps[0], a.steps, err_msg=a.source)
        np.testing.assert_array_equal(flags[0] & twos.CERT_CAPPED, a.capped, err_msg=a.source)
        np.testing.assert_array_equal((flags[0] & twos.CERT_CLIPPED) >> 1, a.clipped, err_msg=a.source)
        assert not np.any(flags & twos.CERT_CONTAINMENT)
        np.testing.assert_allclose(value[0], a.value, rtol=1e-10, atol=0, err_msg=a.source)
        assert np.mean(value[0] == a.value) > 0.5, a.source  # a systematic error would leave none bitwise


@pytest.mark.parametrize("index", range(12))
def test_all_stages_of_an_attempt(ctx, setup, index):
    meshes, attempts = setup
    a = attempts[index]
    print("certify all samples", a.source, flush=True)
    value, steps, flags = _certify(ctx, meshes, a, a.walks)
    for st in a.stages:
        n = st["n"]
        assert int(np.sum(flags[0, :n] & twos.CERT_CAPPED)) == st["path_caps"]
        assert int(np.sum((flags[0, :n] & twos.CERT_CLIPPED) >> 1)) == st["clipped"]
        mean = float(np.sum(value[0, :n].astype(np.longdouble)) / n)
        assert mean == pytest.approx(st["mean"], rel=1e-11, abs=0)
        assert float(np.mean(steps[0, :n])) == pytest.approx(st["mean_steps"], rel=1e-12)


def test_certified_clearance(ctx, setup):
    meshes, attempts = setup
    for a in attempts:
        g = ctx.geometry(twos.MESH, mesh=meshes[a.kind], cavities=a.holes if len(a.holes) else None,
                         cavity_radius=a.radius)
        try:
            got = ctx.cert_distance(g, a.point)[0]
        finally:
            ctx.release_geometry(g)
        assert got == a.stages[0]["center_clearance_lower"], a.source


def test_sample_blocks_compose(ctx, setup):
    meshes, attempts = setup
    a = attempts[3]
    whole = _certify(ctx, meshes, a, 1536)
    tail = _certify(ctx, meshes, a, 512, sample_begin=1024)
    for x, y in zip(whole, tail):
        np.testing.assert_array_equal(x[:, 1024:], y)


def test_wih_twos_validate(setup, golden, tmp_path):
    """wih_twos --mode validate (run_minimum_benefit, minimum_benefit.hpp:
    145-188): the samples file and the staged summary of the reference's
    first attempt, and the fail-closed exit for a site outside the mesh."""
    import json
    import subprocess

    from cli_files import BIN, write_ply, write_points

    _, attempts = setup
    a = attempts[0]
    d = tmp_path
    v, f = CF.mesh(golden, a.kind)
    write_ply(d / "bunny.ply", v, f)
    write_points(d / "points.csv", a.point)
    args = ["--mode", "validate", "--mesh", str(d / "bunny.ply"), "--points", str(d / "points.csv"), "--output",
            str(d / "validation.csv"), "--summary", str(d / "validation.json"), "--walks", "4096", "--seed",
            str(a.seed), "--threads", "8", "--repeats", "1", "--source-mode", "affine", "--source-y-slope", ".65",
            "--epsilon", repr(a.epsilon), "--max-steps", str(a.max_steps), "--validation-radius-cap",
            repr(a.radius_cap), "--validation-value-cap", repr(a.value_cap)]
    print("cli wih_twos --mode validate", a.source, flush=True)
    r = subprocess.run([str(BIN / "wih_twos"), *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    out = np.genfromtxt(d / "validation.csv", delimiter=",", names=True)
    assert len(out) == 4096 and np.all(out["site"] == 0) and np.array_equal(out["sample"], np.arange(4096))
    n = len(a.value)
    np.testing.assert_array_equal(out["steps"][:n], a.steps)
    np.testing.assert_array_equal(out["capped"][:n], a.capped)
    np.testing.assert_array_equal(out["clipped"][:n], a.clipped)
    np.testing.assert_allclose(out["value"][:n], a.value, rtol=1e-10, atol=0)
    s = json.loads((d / "validation.json").read_text())
    assert [st["n"] for st in s["stages"]] == [1024, 4096] and s["containment_failures"] == 0
    for got, ref in zip(s["stages"], a.stages[:2]):
        assert (got["path_caps"], got["clipped"]) == (ref["path_caps"], ref["clipped"])
        assert got["mean"] == pytest.approx(ref["mean"], rel=1e-11)
        assert got["center_clearance_lower"] == ref["center_clearance_lower"]
        assert got["embree_center_distance"] == pytest.approx(ref["embree_center_distance"], rel=1e-6)
    write_points(d / "outside.csv", [[5.0, 5.0, 5.0]])
    r = subprocess.run([str(BIN / "wih_twos"), *args[:5], str(d / "outside.csv"), *args[6:]], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 1 and "candidate classified outside" in r.stderr
