"""The analytic source (source_kind 3, exact_bump_green of study.cpp:56-76)
on an MI355X, through the C-ABI.

- binary64 on the reference's paired mt19937_64 streams on the study's branch
  mesh: every step count equals the oracle's and u, p, score agree to 1e-9
  (the oracle evaluates the bump integral in long double as the reference
  does, the device in binary64);
- binary64 and binary32 Philox walks, persistent schedule, against the oracle
  and the one-step closed value from the centre of the unit ball;
- wih_study --analytic-source 1 writes the oracle's paired values.
"""

import subprocess

import numpy as np
import pytest

import oracle as O
import study_fixtures as SF
from cli_files import BIN, read_csv, write_ply, write_points

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def study(ctx, golden):
    v, f, requests = SF.load(golden)
    config = next(r.config for r in requests if len(r.config.split()) == 19)
    req = next(r for r in requests if "pipeline" in r.source and len(r.holes) == 5)
    return v, f, ctx.mesh(v, f), O.Mesh(v, f), config, req


def _oracle(study, pts, rng, seed, walks):
    v, f, mesh, omesh, config, req = study
    heat, sensor = O.study_terms(config)
    g = O.Geometry(O.MESH, mesh=omesh, cavities=req.holes, cavity_radius=req.cavity_radius)
    return O.estimate(g, pts, [O.side(terms=heat, analytic=True), O.side(terms=sensor, analytic=True)],
                      combine=O.PAIRED, rng=rng, seed=seed, walks=walks, eps=req.epsilon,
                      max_steps=req.max_steps, threads=8)


def _gpu(ctx, study, pts, **kw):
    v, f, mesh, omesh, config, req = study
    heat, sensor = twos.study_terms(config)
    g = ctx.geometry(twos.MESH, mesh=mesh, cavities=req.holes, cavity_radius=req.cavity_radius)
    try:
        return ctx.estimate(g, pts, [twos.side(terms=heat, analytic=True), twos.side(terms=sensor, analytic=True)],
                            combine=twos.PAIRED, eps=req.epsilon, max_steps=req.max_steps, **kw)
    finally:
        ctx.release_geometry(g)


def _compare(res, ref, rtol=1e-9):
    np.testing.assert_array_equal(res.steps, ref["steps"])
    for name in ("u", "v", "score", "same"):
        r = ref[name]
        scale = 1e-3 * np.max(np.abs(r)) + 1e-300
        np.testing.assert_allclose(getattr(res, name), r, rtol=rtol, atol=rtol * scale, err_msg=name)


def test_paired_mt_replay_matches_oracle(ctx, study):
    req = study[5]
    pts = req.points
    print("analytic paired replay", len(pts), "points", flush=True)
    res = _gpu(ctx, study, pts, rng=twos.RNG_MT_PAIRED, precision=64, seed=req.seed, walks=128)
    _compare(res, _oracle(study, pts, O.RNG_MT_PAIRED, req.seed, 128))
    assert np.sum(res.u > 0) >= 10


def test_philox_binary64_matches_oracle(ctx, study):
    req = study[5]
    res = _gpu(ctx, study, req.points, precision=64, seed=17, walks=256)
    _compare(res, _oracle(study, req.points, O.RNG_PHILOX, 17, 256))


@pytest.mark.parametrize("precision", [32, 64])
def test_one_step_from_the_centre(ctx, precision):
    rows = [(twos.TERM_BUMP, 0.2, 0.0, 0.1, 0.3, 0.0, 0.0, 2.0), (twos.TERM_BUMP, -0.3, 0.2, 0.0, 0.25, 0.0, 0.0, 5.0)]
    orows = [(O.TERM_BUMP, *r[1:]) for r in rows]
    expect = sum(O.bump_green_ball([0, 0, 0], 1.0, r[1:4], r[4], r[7]) for r in orows)
    g = ctx.geometry(twos.BALL)
    res = ctx.estimate(g, [[0.0, 0.0, 0.0]], [twos.side(terms=rows, analytic=True)], combine=twos.SQUARE,
                       precision=precision, seed=1, walks=256, eps=1e-6, max_steps=16)
    ctx.release_geometry(g)
    assert res.steps[0] == 256
    assert abs(res.u[0] - expect) <= 1e-6 * expect


def test_wih_study_analytic_source(study, tmp_path):
    v, f, mesh, omesh, config, req = study
    d = tmp_path
    write_ply(d / "branch.ply", v, f)
    (d / "src.cfg").write_text(config)
    write_points(d / "pts.txt", req.points, sep=" ")
    write_points(d / "holes.txt", req.holes, sep=" ")
    args = ["--mesh", str(d / "branch.ply"), "--config", str(d / "src.cfg"), "--points", str(d / "pts.txt"),
            "--cavities", str(d / "holes.txt"), "--cavity-radius", repr(req.cavity_radius), "--output",
            str(d / "field.csv"), "--summary", str(d / "kernel.json"), "--mode", "twos", "--source-mode", "housing",
            "--walks", "64", "--repeats", "1", "--seed", str(req.seed), "--threads", "4", "--max-steps",
            str(req.max_steps), "--epsilon", repr(req.epsilon), "--hot-only", "0", "--fd", "0", "--diff", "0",
            "--analytic-source", "1"]
    print("cli wih_study --analytic-source 1", flush=True)
    r = subprocess.run([str(BIN / "wih_study"), *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    out = read_csv(d / "field.csv")
    ref = _oracle(study, req.points, O.RNG_MT_PAIRED, req.seed, 64)
    inside = np.min(np.linalg.norm(req.points[:, None] - req.holes[None], axis=2), axis=1) <= req.cavity_radius
    live = ~inside
    np.testing.assert_array_equal(out["steps"][live].astype(np.int64), ref["steps"][live])
    for col, name in (("u", "u"), ("p", "v"), ("score", "score"), ("same_index", "same")):
        scale = 1e-3 * np.max(np.abs(ref[name][live])) + 1e-300
        np.testing.assert_allclose(out[col][live], ref[name][live], rtol=1e-9, atol=1e-9 * scale, err_msg=col)
