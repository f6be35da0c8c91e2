"""Differential WoS (twos_differential; evaluate_differential_site,
benchmarks/compiled_bunny/src/wih_twos.cpp:725-928, as wih_study --diff 1
runs it, benchmarks/submission_study/study.cpp:303-316) on an MI355X,
through the C-ABI.

- all fourteen retained --diff 1 requests on the branch mesh: kernel.json's
  request totals (steps, distance queries, parity rays) exactly, the
  derivative columns to 2e-4 of scale and most sites bitwise (near-tie
  closest triangles, and the odd ulp between the device library's and
  glibc's cos / sin / exp, move a few paths by ulps; a WoS step can at most
  double a position perturbation, so they stay small);
- against the oracle per site: every counter (outer and nested steps,
  cavity terminations, seam / segment / unsafe classifications, caps,
  queries, rays) exactly, values as above, also with overlapping cavities;
- a site's result depends on (seed, repeat, site index, sample) only;
- wih_study --diff 1 writes the library's values and the reference's totals;
- the production form (Philox streams, one persistent launch): binary64
  equal to the oracle's Philox form per site (every counter exactly, values
  as above), binary32 within a fraction of an SE of its binary64 rerun, the
  distance-grid walk within its combined SE of the exact walk, and results
  keyed by site index.
"""

import json
import subprocess

import numpy as np
import pytest

import oracle as O
import study_fixtures as SF
from cli_files import BIN, write_ply, write_points

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def diff(ctx, golden):
    v, f, requests = SF.load_diff(golden)
    return v, f, ctx.mesh(v, f), O.Mesh(v, f), requests


def _req(diff, source):
    return next(r for r in diff[4] if r.source == source)


def _gpu(ctx, diff, req, *, holes=None, points=None, walks=None, seed=None, site_offset=0, **kw):
    heat, _ = twos.study_terms(req.config)
    g = ctx.geometry(twos.MESH, mesh=diff[2], cavities=req.holes if holes is None else holes,
                     cavity_radius=req.cavity_radius)
    try:
        return ctx.differential(g, req.points if points is None else points, twos.side(terms=heat),
                                seed=req.seed if seed is None else seed, walks=walks or req.walks,
                                eps=req.epsilon, max_steps=req.max_steps, offset=req.offset,
                                site_offset=site_offset, **kw)
    finally:
        ctx.release_geometry(g)


def _oracle(diff, req, *, holes=None, points=None, walks=None, seed=None, rng=O.RNG_MT_DIFF):
    heat, _ = O.study_terms(req.config)
    g = O.Geometry(O.MESH, mesh=diff[3], cavities=req.holes if holes is None else holes,
                   cavity_radius=req.cavity_radius)
    return O.differential(g, req.points if points is None else points, O.side(terms=heat),
                          seed=req.seed if seed is None else seed, walks=walks or req.walks, eps=req.epsilon,
                          max_steps=req.max_steps, offset=req.offset, threads=O.host_threads(), rng=rng)


PHILOX64 = dict(rng=twos.RNG_PHILOX, precision=64)
PHILOX32 = dict(rng=twos.RNG_PHILOX, precision=32)
PHILOX32_GRID = dict(rng=twos.RNG_PHILOX, precision=32, distance_grid=True)


def _z(a, b, tol_se=0.0):
    """Per (site, cavity, column) z-scores of two results' derivative columns
    in their combined SE, over the entries where that SE is positive, and the
    entries where it is zero (both must then agree exactly)."""
    zs, zero_pairs = [], []
    for col, se in ((0, 3), (1, 4), (2, 5), (6, 7)):
        s = np.hypot(a[..., se], b[..., se]) + tol_se
        ok = s > 0
        zs.append((a[..., col] - b[..., col])[ok] / s[ok])
        zero_pairs.append((a[..., col][~ok], b[..., col][~ok]))
    return np.concatenate(zs), zero_pairs


def _totals(counts):
    c = dict(zip(O.DIFF_COUNTS, counts.sum(axis=0)))
    return {"steps": int(c["outer_steps"] + c["nested_steps"]), "distance_queries": int(c["distance_queries"]),
            "parity_rays": int(c["parity_rays"]), "caps": int(c["limits"])}


SOURCES = [f"validation-local-cavity/raw/gradient-{o}-{k}" for o in ("0.002", "0.004", "0.008") for k in range(4)]
SOURCES += ["validation/raw/gradient-0.002-0", "validation/raw/gradient-0.002-1"]


@pytest.mark.parametrize("source", SOURCES)
def test_requests_match_reference(ctx, diff, source):
    req = _req(diff, source)
    print("differential", source, len(req.points), "sites", len(req.holes), "cavities", flush=True)
    values, counts = _gpu(ctx, diff, req)
    assert _totals(counts) == req.totals
    exact = SF.compare_diff(req, values, rtol=2e-4)
    assert exact.mean() > 0.5, exact.mean()  # a systematic error would leave no site bitwise


def test_counters_match_oracle_per_site(ctx, diff):
    req = _req(diff, "validation-local-cavity/raw/gradient-0.004-1")
    values, counts = _gpu(ctx, diff, req, walks=256, seed=77)
    ref_values, ref_counts = _oracle(diff, req, walks=256, seed=77)
    np.testing.assert_array_equal(counts, ref_counts)
    ref = SF.DiffRequest(**{**req.__dict__, "values": ref_values})
    exact = SF.compare_diff(ref, values, rtol=2e-4)
    assert exact.mean() > 0.5, exact.mean()  # a systematic error would leave no site bitwise
    assert np.count_nonzero(values) > 0


def test_overlapping_cavities_match_oracle(ctx, diff):
    """Two cavities closer than twice the radius and a third apart: seam and
    segment classifications, per site, as the oracle makes them."""
    req = _req(diff, "validation-local-cavity/raw/gradient-0.002-0")
    c = req.holes[0]
    holes = np.array([c, c + [0.6 * req.cavity_radius, 0.0, 0.0], c + [0.0, 3.0 * req.cavity_radius, 0.0]])
    pts = req.points[:32]
    values, counts = _gpu(ctx, diff, req, holes=holes, points=pts, walks=128, seed=3)
    ref_values, ref_counts = _oracle(diff, req, holes=holes, points=pts, walks=128, seed=3)
    np.testing.assert_array_equal(counts, ref_counts)
    tot = dict(zip(O.DIFF_COUNTS, counts.sum(axis=0)))
    assert tot["near_seam"] > 0 and tot["unsafe"] >= tot["near_seam"]
    ref = SF.DiffRequest(**{**req.__dict__, "points": pts, "holes": holes, "values": ref_values})
    SF.compare_diff(ref, values, rtol=2e-4)


def test_results_depend_on_site_index_only(ctx, diff):
    req = _req(diff, "validation-local-cavity/raw/gradient-0.002-1")
    full, fc = _gpu(ctx, diff, req, walks=64)
    head, hc = _gpu(ctx, diff, req, points=req.points[:20], walks=64)
    np.testing.assert_array_equal(head, full[:20])
    np.testing.assert_array_equal(hc, fc[:20])
    mid, _ = _gpu(ctx, diff, req, points=req.points[40:50], walks=64, site_offset=40)
    np.testing.assert_array_equal(mid, full[40:50])


def test_philox_binary64_matches_oracle(ctx, diff):
    """The production streams in binary64 without the grid: every per-site
    counter equal to the oracle's Philox form, the derivative columns to 2e-4
    of scale with most sites bitwise (the device library's cos / sin may
    differ from glibc's by an ulp)."""
    req = _req(diff, "validation-local-cavity/raw/gradient-0.004-1")
    print("differential philox binary64", req.source, flush=True)
    values, counts = _gpu(ctx, diff, req, walks=256, seed=77, **PHILOX64)
    ref_values, ref_counts = _oracle(diff, req, walks=256, seed=77, rng=O.RNG_PHILOX)
    np.testing.assert_array_equal(counts, ref_counts)
    ref = SF.DiffRequest(**{**req.__dict__, "values": ref_values})
    exact = SF.compare_diff(ref, values, rtol=2e-4)
    assert exact.mean() > 0.5, exact.mean()
    assert np.count_nonzero(values) > 0


def test_philox_binary64_overlapping_cavities_match_oracle(ctx, diff):
    req = _req(diff, "validation-local-cavity/raw/gradient-0.002-0")
    c = req.holes[0]
    holes = np.array([c, c + [0.6 * req.cavity_radius, 0.0, 0.0], c + [0.0, 3.0 * req.cavity_radius, 0.0]])
    pts = req.points[:32]
    print("differential philox binary64 overlapping cavities", flush=True)
    values, counts = _gpu(ctx, diff, req, holes=holes, points=pts, walks=128, seed=3, **PHILOX64)
    ref_values, ref_counts = _oracle(diff, req, holes=holes, points=pts, walks=128, seed=3, rng=O.RNG_PHILOX)
    np.testing.assert_array_equal(counts, ref_counts)
    tot = dict(zip(O.DIFF_COUNTS, counts.sum(axis=0)))
    assert tot["near_seam"] > 0 and tot["unsafe"] >= tot["near_seam"]
    ref = SF.DiffRequest(**{**req.__dict__, "points": pts, "holes": holes, "values": ref_values})
    SF.compare_diff(ref, values, rtol=2e-4)


def test_philox_binary32_shadows_binary64(ctx, diff):
    """binary32 walks draw the same Philox numbers as their binary64 rerun, so
    most paths take the same steps: derivatives within a fraction of their SE
    and step totals within 1%."""
    req = _req(diff, "validation-local-cavity/raw/gradient-0.004-0")
    print("differential philox binary32 vs binary64", flush=True)
    v64, c64 = _gpu(ctx, diff, req, walks=1024, seed=5, **PHILOX64)
    v32, c32 = _gpu(ctx, diff, req, walks=1024, seed=5, **PHILOX32)
    z, zero = _z(v32, v64)
    assert len(z) > 300
    # A path that ends differently moves its site by about one sample / N.
    assert np.mean(np.abs(z) > 1.0) <= 0.02 and np.max(np.abs(z)) < 3.0, np.sort(np.abs(z))[-4:]
    for a, b in zero:
        np.testing.assert_array_equal(a, b)
    for k in (0, 1, 2, 7):
        assert abs(c32[:, k].sum() - c64[:, k].sum()) <= 0.01 * c64[:, k].sum(), (O.DIFF_COUNTS[k], c32[:, k].sum(),
                                                                                   c64[:, k].sum())


def test_philox_distance_grid_against_exact(ctx, diff):
    """Mesh steps far from the mesh take the distance grid's conservative
    bound as their radius: other paths, the same derivatives -- within their
    combined SE of the exact binary64 walk -- fewer exact queries, and the
    same rate of cavity terminations."""
    req = _req(diff, "validation-local-cavity/raw/gradient-0.004-0")
    print("differential philox binary32 grid vs exact binary64", flush=True)
    ex, ce = _gpu(ctx, diff, req, walks=2048, seed=9, **PHILOX64)
    gr, cg = _gpu(ctx, diff, req, walks=2048, seed=9, **PHILOX32_GRID)
    assert not np.array_equal(cg[:, 0], ce[:, 0])  # the grid was used
    z, _ = _z(gr, ex)
    assert len(z) > 300
    assert np.sum(np.abs(z) > 3.5) <= 2 and abs(np.mean(z)) < 0.2, (np.sort(np.abs(z))[-4:], np.mean(z))
    te, tg = ce.sum(axis=0), cg.sum(axis=0)
    assert abs(tg[2] - te[2]) <= 0.05 * te[2], (tg[2], te[2])  # cavity terminations


def test_philox_results_depend_on_site_index_only(ctx, diff):
    req = _req(diff, "validation-local-cavity/raw/gradient-0.002-1")
    full, fc = _gpu(ctx, diff, req, walks=64, **PHILOX32_GRID)
    head, hc = _gpu(ctx, diff, req, points=req.points[:20], walks=64, **PHILOX32_GRID)
    np.testing.assert_array_equal(head, full[:20])
    np.testing.assert_array_equal(hc, fc[:20])
    mid, _ = _gpu(ctx, diff, req, points=req.points[40:50], walks=64, site_offset=40, **PHILOX32_GRID)
    np.testing.assert_array_equal(mid, full[40:50])
    again, _ = _gpu(ctx, diff, req, walks=64, **PHILOX32_GRID)
    np.testing.assert_array_equal(again, full)


def test_rejects_non_mesh_geometry(ctx):
    g = ctx.geometry(twos.BALL, cavities=[[0.0, 0.0, 0.0]], cavity_radius=0.1)
    try:
        with pytest.raises(twos.TwosError):
            ctx.differential(g, [[0.5, 0.0, 0.0]], twos.side(), walks=4)
    finally:
        ctx.release_geometry(g)


def test_wih_study_diff(ctx, diff, tmp_path):
    req = _req(diff, "validation-local-cavity/raw/gradient-0.008-0")
    v, f = diff[0], diff[1]
    d = tmp_path
    write_ply(d / "branch.ply", v, f)
    (d / "branch.txt").write_text(req.config)
    write_points(d / "pts.txt", req.points, sep=" ")
    write_points(d / "holes.txt", req.holes, sep=" ")
    args = ["--mesh", str(d / "branch.ply"), "--config", str(d / "branch.txt"), "--points", str(d / "pts.txt"),
            "--cavities", str(d / "holes.txt"), "--cavity-radius", repr(req.cavity_radius), "--output",
            str(d / "field.csv"), "--summary", str(d / "kernel.json"), "--mode", "twos", "--source-mode", "housing",
            "--walks", str(req.walks), "--repeats", "1", "--seed", str(req.seed), "--threads", "4",
            "--max-steps", str(req.max_steps), "--epsilon", repr(req.epsilon), "--hot-only", "0", "--fd", "0",
            "--diff", "1", "--grid", "65,47,27", "--mg-cycles", "120", "--mg-tolerance", "1e-7",
            "--diff-offset", repr(req.offset)]
    print("cli wih_study --diff 1", req.source, flush=True)
    r = subprocess.run([str(BIN / "wih_study"), *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    out = np.genfromtxt(d / "field.csv", delimiter=",", names=True)
    assert out.dtype.names[:4] == ("site", "x", "y", "z")
    got = np.stack([out[n] for n in out.dtype.names[4:]], axis=1).reshape(req.values.shape)
    values, _ = _gpu(ctx, diff, req)
    np.testing.assert_array_equal(got, values)
    summary = json.loads((d / "kernel.json").read_text())
    assert {k: summary[k] for k in req.totals} == req.totals


def test_wih_twos_diffwos(ctx, diff, tmp_path):
    """wih_twos --mode diffwos (run_differential_wos, wih_twos.cpp:1604-1705):
    the affine source of --source-mode affine, repeat 0's field and counters
    over two repeats, equal to the library's call."""
    req = _req(diff, "validation-local-cavity/raw/gradient-0.004-2")
    d = tmp_path
    write_ply(d / "branch.ply", diff[0], diff[1])
    write_points(d / "pts.csv", req.points)
    write_points(d / "holes.csv", req.holes)
    args = ["--mesh", str(d / "branch.ply"), "--points", str(d / "pts.csv"), "--cavities", str(d / "holes.csv"),
            "--cavity-radius", repr(req.cavity_radius), "--output", str(d / "field.csv"), "--summary",
            str(d / "summary.json"), "--mode", "diffwos", "--source-mode", "affine", "--walks", "256",
            "--repeats", "2", "--seed", "91", "--max-steps", str(req.max_steps), "--epsilon", repr(req.epsilon),
            "--diff-offset", repr(req.offset)]
    print("cli wih_twos --mode diffwos", flush=True)
    r = subprocess.run([str(BIN / "wih_twos"), *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    out = np.genfromtxt(d / "field.csv", delimiter=",", names=True)
    got = np.stack([out[n] for n in out.dtype.names[4:]], axis=1).reshape(len(req.points), len(req.holes), 8)
    g = ctx.geometry(twos.MESH, mesh=diff[2], cavities=req.holes, cavity_radius=req.cavity_radius)
    try:
        values, counts = ctx.differential(g, req.points, twos.reference_affine_side(), seed=91, walks=256,
                                          eps=req.epsilon, max_steps=req.max_steps, offset=req.offset)
    finally:
        ctx.release_geometry(g)
    np.testing.assert_array_equal(got, values)
    assert np.count_nonzero(values) > 0
    s = json.loads((d / "summary.json").read_text())
    c = counts.sum(axis=0)
    assert s["mode"] == "diffwos" and len(s["timings_seconds"]) == 2
    assert (s["distance_queries"], s["parity_rays"]) == (c[7], c[8])
    assert s["outer_mean_steps"] == pytest.approx(c[0] / (256 * len(req.points)), rel=1e-12)
