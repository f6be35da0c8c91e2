"""Sampled volume sources on an MI355X, through the C-ABI.

- The paired estimator's retained outputs (benchmarks/submission_study,
  paired_site, study.cpp:80-120): binary64 walks on the reference's
  mt19937_64 streams with the heat and sensor bumps sampled at a shared
  Beta(2,2) Green point reproduce every step and cap count exactly and the
  values to 1e-6 (see study_fixtures.compare for why not every site is
  bitwise).
- Green-point moments in closed form, binary64 and binary32 walks.
- Philox per-path samples of a sampled source against the oracle.
"""

import math

import numpy as np
import pytest

import oracle as O
import study_fixtures as SF

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
    return ctx.mesh(v, f), requests


def test_study_requests_on_gpu(ctx, study):
    mesh, requests = study
    exact = []
    for req in requests:
        print("study request", req.source, len(req.points), "sites", req.walks, "walks", flush=True)
        heat, sensor = twos.study_terms(req.config)
        g = ctx.geometry(twos.MESH, mesh=mesh, cavities=req.holes if len(req.holes) else None,
                         cavity_radius=req.cavity_radius)
        sides = [twos.side(terms=heat), twos.side() if req.hot else twos.side(terms=sensor)]
        r = ctx.estimate(g, req.points, sides, combine=twos.PAIRED, rng=twos.RNG_MT_PAIRED, precision=64,
                         seed=req.seed, walks=req.walks, eps=req.epsilon, max_steps=req.max_steps)
        if req.hot:
            zero = np.zeros_like(r.u)
            exact.append(SF.compare(req, r.u, zero, r.u, zero, r.steps, r.caps, rtol=1e-6))
        else:
            exact.append(SF.compare(req, r.u, r.v, r.score, r.same, r.steps, r.caps, rtol=1e-6))
        ctx.release_geometry(g)
    assert np.concatenate(exact).mean() > 0.5


def _ball(ctx):
    return ctx.geometry(twos.BALL)


@pytest.mark.parametrize("precision,rng", [(64, twos.RNG_PHILOX), (32, twos.RNG_PHILOX), (64, twos.RNG_MT_SITE)])
def test_green_point_moments_on_gpu(ctx, precision, rng):
    """One step from the centre of the unit ball (eps = 0.99): the value is
    f(a) / 6 at one Green point; E f for the bump w (1 - |a - c|^2 / R^2)^2 is
    w (1 - 2 (3/10 + |c|^2) / R^2 + (1/7 + |c|^2 + |c|^4) / R^4)."""
    c = np.array([0.3, -0.2, 0.1])
    w, R = 1.5, 2.0
    bump = twos.side(terms=[(twos.TERM_BUMP, *c, R, 0.0, 0.0, w)])
    n = 1 << 18
    r = ctx.estimate(_ball(ctx), [[0.0, 0.0, 0.0]], [bump], combine=twos.SQUARE, rng=rng, precision=precision,
                     schedule=twos.PLAIN if rng != twos.RNG_PHILOX else twos.PERSISTENT, seed=4, walks=n,
                     eps=0.99, max_steps=8, dump=True)
    assert r.steps[0] == n
    vals = r.a[0]
    c2 = float(c @ c)
    expect = w * (1 - 2 * (0.3 + c2) / R**2 + (1 / 7 + c2 + c2**2) / R**4) / 6
    se = vals.std() / math.sqrt(n)
    assert abs(vals.mean() - expect) < 4 * se + (2e-6 if precision == 32 else 0.0)


@pytest.mark.parametrize("schedule", [twos.PLAIN, twos.PERSISTENT])
def test_sampled_source_philox_matches_oracle(ctx, schedule):
    """Paired binary64 walks in a box with a Gaussian-chip source (the
    housing source of wih_twos.cpp:553-566, recentred) and a bump: per-path
    samples equal the oracle's to 1e-9 (glibc and the device library may
    round sin/cos/exp by an ulp), step counts exactly."""
    off, chips = twos.housing_chip_terms()
    chips = [(t[0], t[1] + 0.5, t[2] + 0.5, t[3] + 0.5, *t[4:]) for t in chips]
    bump = [(twos.TERM_BUMP, 0.5, 0.5, 0.5, 0.4, 0.0, 0.0, 3.0)]
    sites = np.array([[0.5, 0.5, 0.5], [0.2, 0.7, 0.4], [0.8, 0.3, 0.6], [0.1, 0.1, 0.9]])
    g = ctx.geometry(twos.BOX)
    res = ctx.estimate(g, sites, [twos.side(terms=chips, offset=off), twos.side(terms=bump)],
                       combine=twos.PAIRED, precision=64, schedule=schedule, seed=11, walks=512, eps=1e-3,
                       max_steps=4096, dump=True)
    ref = O.estimate(O.Geometry(O.BOX), sites, [O.side(terms=chips, offset=off), O.side(terms=bump)],
                     combine=O.PAIRED, rng=O.RNG_PHILOX, seed=11, walks=512, eps=1e-3, max_steps=4096, dump=True)
    np.testing.assert_array_equal(res.walk_steps, ref["walk_steps"])
    np.testing.assert_allclose(res.a, ref["a"], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(res.b, ref["b"], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(res.score, ref["score"], rtol=1e-9, atol=1e-13)


def test_sampled_source_binary32_shadows_binary64(ctx):
    """binary32 walks draw the same Philox numbers (24-bit uniforms for the
    Beta radius in both precisions), so their site means differ from the
    binary64 rerun far less than the sampling error."""
    bump = [(twos.TERM_BUMP, 0.5, 0.5, 0.5, 0.4, 0.0, 0.0, 3.0)]
    sites = np.array([[0.5, 0.5, 0.5], [0.3, 0.6, 0.4]])
    g = ctx.geometry(twos.BOX)
    kw = dict(combine=twos.SQUARE, seed=5, walks=4096, eps=1e-3, max_steps=4096, dump=True)
    r32 = ctx.estimate(g, sites, [twos.side(terms=bump)], precision=32, **kw)
    r64 = ctx.estimate(g, sites, [twos.side(terms=bump)], precision=64, **kw)
    se = r64.a.std(axis=1) / math.sqrt(r64.a.shape[1])
    assert np.max(np.abs(r32.u - r64.u) / se) < 0.5


def test_housing_fields_on_gpu(ctx, golden):
    """The engineering-housing scout/validation fields (evaluate_site with the
    housing chip source at a Green point, binary32 query radius): binary64
    walks on the reference's mt19937_64 streams reproduce every step count
    and u / score / score_se to 1e-6 (glibc and device sin/cos/exp may differ
    by an ulp, which a path carries forward)."""
    v, f, requests = SF.load_housing(golden)
    mesh = ctx.mesh(v, f)
    off, chips = twos.housing_chip_terms()
    for req in requests:
        print("housing request", req.tag, len(req.points), "sites", req.walks, "walks", flush=True)
        g = ctx.geometry(twos.MESH, mesh=mesh, cavities=req.cavities if len(req.cavities) else None,
                         cavity_radius=req.cavity_radius, mesh_radius_f32=True)
        r = ctx.estimate(g, req.points, [twos.side(terms=chips, offset=off)], combine=twos.SQUARE,
                         rng=twos.RNG_MT_SITE, precision=64, seed=req.seed, walks=req.walks, eps=req.epsilon,
                         max_steps=req.max_steps)
        SF.compare_housing(req, r.u, r.score, r.score_se, r.steps, r.caps, rtol=1e-6)
        ctx.release_geometry(g)
    ctx.release_mesh(mesh)
