"""The coupled finite-cavity difference (TWOS_COUPLED, study.cpp --fd) on an
MI355X, through the C-ABI.

- binary64 on the reference's --fd streams (RNG_MT_FD) on the study's branch
  mesh with its heat bumps, sensor and cavities: every step and cap count
  equals the oracle's, the score and its SE agree to 1e-9 (the device and
  glibc cos/sin/cbrt may differ in the last bit);
- binary64 Philox, plain and persistent schedules, against the oracle;
- the concentric unit-ball closed form in binary32 and binary64, and a
  cavity no path reaches giving exact zeros;
- a COUPLED request gives the same results alone and mixed with a CROSS
  request in one server batch;
- wih_study --fd 1 writes the oracle's values (u and p zero, score, SE in
  same_index).
"""

import json
import subprocess

import numpy as np
import pytest
from scipy import integrate

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


class Case:
    """The branch mesh, the current source configuration (heat bumps and the
    sensor's own heating, so the sensor region is warm), the cavities and
    seed of a pilot design request, and candidate points near the sensor
    (where a cavity changes J): the request's nearest points and points
    0.12 from the sensor centre."""

    def __init__(self, ctx, golden):
        v, f, requests = SF.load(golden)
        self.config = next(r.config for r in requests if len(r.config.split()) == 19)
        req = next(r for r in requests if "pipeline" in r.source and len(r.holes) == 5)
        self.vertices, self.faces = v, f
        self.mesh, self.omesh = ctx.mesh(v, f), O.Mesh(v, f)
        self.holes, self.radius, self.seed = req.holes, req.cavity_radius, req.seed
        self.eps, self.max_steps = req.epsilon, req.max_steps
        sc = np.array([float(x) for x in self.config.split()[:3]])
        near = req.points[np.argsort(np.linalg.norm(req.points - sc, axis=1))[:6]]
        d = np.random.default_rng(0).normal(size=(14, 3))
        self.points = np.vstack([near, sc + 0.12 * d / np.linalg.norm(d, axis=1)[:, None]])

    def gpu(self, ctx):
        heat, sensor = twos.study_terms(self.config)
        g = ctx.geometry(twos.MESH, mesh=self.mesh, cavities=self.holes, cavity_radius=self.radius)
        return g, [twos.side(terms=heat), twos.side(terms=sensor)]

    def oracle(self, pts, rng, seed, walks):
        heat, sensor = O.study_terms(self.config)
        g = O.Geometry(O.MESH, mesh=self.omesh, cavities=self.holes, cavity_radius=self.radius)
        return O.estimate(g, pts, [O.side(terms=heat), O.side(terms=sensor)], combine=O.COUPLED, rng=rng,
                          seed=seed, walks=walks, eps=self.eps, max_steps=self.max_steps, threads=8)


@pytest.fixture(scope="module")
def branch(ctx, golden):
    return Case(ctx, golden)


def _compare(res, ref, rtol=1e-9):
    np.testing.assert_array_equal(res.steps, ref["steps"])
    np.testing.assert_array_equal(res.caps, ref["caps"])
    scale = 1e-3 * np.max(np.abs(ref["score"])) + 1e-300
    np.testing.assert_allclose(res.score, ref["score"], rtol=rtol, atol=rtol * scale)
    np.testing.assert_allclose(res.score_se, ref["score_se"], rtol=1e-7, atol=rtol * scale)


def test_fd_streams_match_oracle_on_the_branch_mesh(ctx, branch):
    c = branch
    pts = c.points[:12]
    print("fd replay", len(pts), "points", len(c.holes), "holes", flush=True)
    g, sides = c.gpu(ctx)
    res = ctx.estimate(g, pts, sides, combine=twos.COUPLED, rng=twos.RNG_MT_FD, precision=64, seed=c.seed,
                       walks=64, eps=c.eps, max_steps=c.max_steps)
    ctx.release_geometry(g)
    ref = c.oracle(pts, O.RNG_MT_FD, c.seed, 64)
    _compare(res, ref)
    assert np.sum(res.score != 0.0) >= 6


@pytest.mark.parametrize("schedule", ["plain", "persistent"])
def test_philox_binary64_matches_oracle(ctx, branch, schedule):
    c = branch
    pts = c.points[4:20]
    g, sides = c.gpu(ctx)
    res = ctx.estimate(g, pts, sides, combine=twos.COUPLED, precision=64, seed=9, walks=96, eps=c.eps,
                       max_steps=c.max_steps, schedule=twos.PLAIN if schedule == "plain" else twos.PERSISTENT)
    ctx.release_geometry(g)
    ref = c.oracle(pts, O.RNG_PHILOX, 9, 96)
    _compare(res, ref)
    assert np.sum(res.score != 0.0) >= 6


def _concentric_exact(rc, rs):
    b = -rc * (1 + rc) / 6
    a = 1 / 6 - b
    u0 = lambda p: (1 - p * p) / 6
    u1 = lambda p: 0.0 if p < rc else -p * p / 6 + a + b / p
    q = lambda p: (1 - (p / rs) ** 2) ** 2 * 105 / (32 * np.pi * rs**3)
    diff = integrate.quad(lambda p: 4 * np.pi * p * p * q(p) * (u0(p) - u1(p)), 0, rs, points=[rc], limit=200)[0]
    return diff / (4 * np.pi * rc)


@pytest.mark.parametrize("precision", [32, 64])
def test_concentric_closed_form(ctx, precision):
    rc, rs = 0.2, 0.6
    exact = _concentric_exact(rc, rs)
    g = ctx.geometry(twos.BALL, cavity_radius=rc)
    const = twos.side(source=(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    res = ctx.estimate(g, [[0.0, 0.0, 0.0]], [const, twos.probe_side((0.0, 0.0, 0.0), rs)], combine=twos.COUPLED,
                       precision=precision, seed=21, walks=1 << 18, eps=1e-5, max_steps=8192)
    ctx.release_geometry(g)
    score, se = res.score[0], res.score_se[0]
    print("concentric", precision, score, exact, se, flush=True)
    assert se < 0.01 * exact
    assert abs(score - exact) < 4 * se


@pytest.mark.parametrize("precision", [32, 64])
def test_cavity_out_of_reach_is_exactly_zero(ctx, precision):
    g = ctx.geometry(twos.BALL, cavity_radius=0.1)
    const = twos.side(source=(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    res = ctx.estimate(g, [[5.0, 0.0, 0.0], [0.0, -4.0, 0.5]], [const, twos.probe_side((0.1, 0.2, 0.0), 0.3)],
                       combine=twos.COUPLED, precision=precision, seed=3, walks=4096, eps=1e-4, max_steps=4096)
    ctx.release_geometry(g)
    np.testing.assert_array_equal(res.score, [0.0, 0.0])
    np.testing.assert_array_equal(res.score_se, [0.0, 0.0])
    assert np.all(res.steps % 2 == 0)


def test_same_alone_and_in_a_mix(ctx, branch):
    c = branch
    g, sides = c.gpu(ctx)
    kw = dict(combine=twos.COUPLED, precision=32, seed=5, walks=128, walks_max=2048, ci_rel=0.2, eps=c.eps,
              max_steps=c.max_steps)
    alone = ctx.estimate(g, c.points, sides, **kw)
    srv = ctx.server(max_sites=4096, max_walks_per_round=1 << 20)
    other = srv.submit(g, c.points[::-1], sides, combine=twos.CROSS, seed=6, walks=64, eps=c.eps,
                       max_steps=c.max_steps)
    mine = srv.submit(g, c.points, sides, **kw)
    done = {}
    while srv.pending() != (0, 0):
        done.update(srv.step())
    srv.close()
    ctx.release_geometry(g)
    assert other in done
    mixed = done[mine]
    np.testing.assert_array_equal(mixed.score, alone.score)
    np.testing.assert_array_equal(mixed.score_se, alone.score_se)
    np.testing.assert_array_equal(mixed.steps, alone.steps)
    np.testing.assert_array_equal(mixed.walks, alone.walks)


def test_wih_study_fd(branch, tmp_path):
    c = branch
    d = tmp_path
    write_ply(d / "branch.ply", c.vertices, c.faces)
    (d / "src.cfg").write_text(c.config)
    pts = c.points[:10]
    write_points(d / "pts.txt", pts, sep=" ")
    write_points(d / "holes.txt", c.holes, sep=" ")
    args = ["--mesh", str(d / "branch.ply"), "--config", str(d / "src.cfg"), "--points", str(d / "pts.txt"),
            "--cavities", str(d / "holes.txt"), "--cavity-radius", repr(c.radius), "--output",
            str(d / "field.csv"), "--summary", str(d / "kernel.json"), "--mode", "twos", "--source-mode", "housing",
            "--walks", "48", "--repeats", "1", "--seed", "731201", "--threads", "4", "--max-steps",
            str(c.max_steps), "--epsilon", repr(c.eps), "--hot-only", "0", "--fd", "1", "--diff", "0"]
    print("cli wih_study --fd 1", flush=True)
    r = subprocess.run([str(BIN / "wih_study"), *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    out = read_csv(d / "field.csv")
    ref = c.oracle(pts, O.RNG_MT_FD, 731201, 48)
    np.testing.assert_array_equal(out["u"], 0.0)
    np.testing.assert_array_equal(out["p"], 0.0)
    np.testing.assert_array_equal(out["steps"].astype(np.int64), ref["steps"])
    scale = 1e-3 * np.max(np.abs(ref["score"])) + 1e-300
    np.testing.assert_allclose(out["score"], ref["score"], rtol=1e-9, atol=1e-9 * scale)
    np.testing.assert_allclose(out["same_index"], ref["score_se"], rtol=1e-7, atol=1e-9 * scale)
    s = json.loads((d / "kernel.json").read_text())
    assert s["steps"] == int(ref["steps"].sum()) and s["caps"] == 0
