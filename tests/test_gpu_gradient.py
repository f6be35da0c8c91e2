"""The first-step gradient channel on an MI355X (TWOS_FLAG_GRADIENT).

- binary64 Philox walks: the GPU's grad_u, grad_v, grad_dot and its SE equal
  the oracle's (same streams, same first steps; the 28 sums are folded in a
  different order, so to 1e-9), plain and persistent schedules; the same on
  the reference's mt19937_64 streams;
- configuration 1 in binary32: grad g_u . grad g_v = 2xz - 2y + 0.5x within
  the estimator's error;
- on a mesh, and in a server mix: the value channel of a request does not
  change with other requests' gradient settings.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402

GU = [0, 0, 0, 0.5, 1, -1, 0, 0, 0, 0]
GV = [0, 0, 1, 0, 0, 0, 0, 0, 1, 0]


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


def _sites(n, seed=0):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d * (0.8 * rng.uniform(0, 1, n) ** (1 / 3))[:, None]


def _compare(res, ref):
    np.testing.assert_array_equal(res.steps, ref["steps"])
    np.testing.assert_allclose(res.grad_u, ref["grad_u"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(res.grad_v, ref["grad_v"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(res.grad_dot, ref["grad_dot"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(res.grad_dot_se, ref["grad_dot_se"], rtol=1e-7)
    np.testing.assert_allclose(res.score, ref["score"], rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("schedule", ["plain", "persistent"])
@pytest.mark.parametrize("combine", ["CROSS", "PAIRED"])
def test_gradient_binary64_matches_oracle(ctx, schedule, combine):
    z = _sites(12)
    cmb = getattr(twos, combine)
    src = (1.0, 0.2, -0.3, 0.1, 0.0, 0.5, 0.0)  # an affine source exercises the r0^2/10 term
    sides_g = [twos.side(poly=GU, poly_classes=(0,), source=src), twos.side(poly=GV, poly_classes=(0,))]
    sides_o = [O.side(poly=GU, poly_classes=(0,), source=src), O.side(poly=GV, poly_classes=(0,))]
    res = ctx.estimate(ctx.geometry(twos.BALL), z, sides_g, combine=cmb, precision=64, seed=3,
                       schedule=twos.PLAIN if schedule == "plain" else twos.PERSISTENT, walks=512, eps=1e-3,
                       max_steps=4096, gradient=True)
    ref = O.estimate(O.Geometry(O.BALL), z, sides_o, combine=getattr(O, combine), rng=O.RNG_PHILOX, seed=3,
                     walks=512, eps=1e-3, max_steps=4096, gradient=True)
    _compare(res, ref)


def test_gradient_mt_replay_matches_oracle(ctx):
    z = _sites(6, seed=4)
    side_g = [twos.side(poly=GU, poly_classes=(0,))]
    side_o = [O.side(poly=GU, poly_classes=(0,))]
    res = ctx.estimate(ctx.geometry(twos.BALL), z, side_g, combine=twos.SQUARE, rng=twos.RNG_MT_SITE,
                       precision=64, seed=9, walks=256, eps=1e-3, max_steps=4096, gradient=True)
    ref = O.estimate(O.Geometry(O.BALL), z, side_o, combine=O.SQUARE, rng=O.RNG_MT_SITE, seed=9, walks=256,
                     eps=1e-3, max_steps=4096, gradient=True)
    _compare(res, ref)


def test_gradient_c1_binary32(ctx):
    z = _sites(64)
    res = ctx.estimate(ctx.geometry(twos.BALL), z, [twos.side(poly=GU, poly_classes=(0,)),
                                                    twos.side(poly=GV, poly_classes=(0,))],
                       combine=twos.CROSS, precision=32, seed=5, walks=16384, eps=1e-3, max_steps=4096,
                       gradient=True)
    x, y, w = z.T
    err = (res.grad_dot - (2 * x * w - 2 * y + 0.5 * x)) / res.grad_dot_se
    assert np.mean(np.abs(err) < 3) >= 0.95 and np.max(np.abs(err)) < 5, err
    assert abs(np.mean(err)) < 0.4


def test_gradient_on_mesh_and_in_a_mix(ctx):
    ico = W.sh_icosphere(n_sites=256)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    g = ctx.geometry(twos.MESH, mesh=mesh)
    sides = W.sides(ico, twos.side)
    kw = dict(combine=twos.CROSS, precision=32, walks=64, walks_max=512, ci_rel=0.05, eps=1e-4, max_steps=8192,
              distance_grid=True)
    grad = ctx.estimate(g, ico.sites[:64], sides, seed=1, gradient=True, **kw)
    assert np.all(np.isfinite(grad.grad_dot)) and np.all(np.isfinite(grad.grad_dot_se))
    plain = ctx.estimate(g, ico.sites[:64], sides, seed=1, **kw)
    assert plain.grad_dot is None  # not requested
    for k in ("u", "v", "score", "score_se", "walks", "steps"):
        np.testing.assert_array_equal(getattr(grad, k), getattr(plain, k), err_msg=k)
    srv = ctx.server(4096, 1 << 20)
    a = srv.submit(g, ico.sites[:64], sides, seed=1, **kw)
    srv.submit(g, ico.sites[64:160], sides, seed=2, **{**kw, "distance_grid": True})
    done = {}
    while srv.pending() != (0, 0):
        done.update(srv.step())
    srv.close()
    for k in ("u", "v", "score", "score_se", "walks", "steps"):
        np.testing.assert_array_equal(getattr(done[a], k), getattr(plain, k), err_msg=k)


def test_gradient_rejects_sampled_sources(ctx):
    bump = twos.side(terms=[(twos.TERM_BUMP, 0, 0, 0, 0.5, 0, 0, 1.0)])
    with pytest.raises(twos.InvalidArgument):
        ctx.estimate(ctx.geometry(twos.BALL), [[0.1, 0, 0]], [bump], gradient=True, walks=16)
