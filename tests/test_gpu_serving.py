"""Continuous-batching server on the GPU: a request's results are the same
alone and mixed with others, admission waits for free slots, and insertion
validations see the cavity."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def env():
    ctx = twos.Context()
    ico = W.sh_icosphere(n_sites=2048)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    g_mesh = ctx.geometry(twos.MESH, mesh=mesh)
    ball = W.unit_ball(n_sites=64)
    g_ball = ctx.geometry(twos.BALL)
    yield ctx, ico, g_mesh, ball, g_ball
    ctx.close()


def _drain(srv, limit=10000):
    out = {}
    for _ in range(limit):
        out.update(srv.step())
        if srv.pending() == (0, 0):
            return out
    raise AssertionError("server did not drain")


KEYS = ("u", "v", "score", "score_se", "walks", "steps", "caps", "status")


def test_same_alone_and_in_the_mix(env):
    ctx, ico, g_mesh, ball, g_ball = env
    kw = dict(combine=twos.CROSS, precision=32, walks=64, walks_max=2048, ci_rel=0.04, eps=1e-4, max_steps=8192,
              distance_grid=True)
    sides_m = W.sides(ico, twos.side)
    sides_b = W.sides(ball, twos.side)
    a_sites = ico.sites[:96]

    alone_srv = ctx.server(4096, 1 << 22)
    ra = alone_srv.submit(g_mesh, a_sites, sides_m, seed=101, **kw)
    alone = _drain(alone_srv)[ra]
    alone_srv.close()

    mix_srv = ctx.server(4096, 1 << 20)
    mix_srv.submit(g_ball, ball.sites, sides_b, seed=7, **kw)
    mix_srv.step()  # another request is already in flight when A arrives
    ra2 = mix_srv.submit(g_mesh, a_sites, sides_m, seed=101, **kw)
    mix_srv.submit(g_mesh, ico.sites[500:900], sides_m, seed=8, **kw)
    mixed = _drain(mix_srv)[ra2]
    mix_srv.close()

    direct = ctx.estimate(g_mesh, a_sites, sides_m, seed=101, **kw)
    for k in KEYS:
        np.testing.assert_array_equal(getattr(alone, k), getattr(mixed, k), err_msg=k)
        np.testing.assert_array_equal(getattr(alone, k), getattr(direct, k), err_msg=k)


def test_grid_choice_is_per_request_in_a_shared_launch(env):
    """Requests of one class (binary32, CROSS) share each persistent launch;
    the distance-grid choice and its threshold stay per request: an exact
    request mixed with grid requests equals the exact request alone, and a
    grid request with another threshold equals itself alone."""
    ctx, ico, g_mesh, _, _ = env
    kw = dict(combine=twos.CROSS, precision=32, walks=64, walks_max=512, ci_rel=0.04, eps=1e-4, max_steps=8192)
    sides_m = W.sides(ico, twos.side)
    exact_sites, grid_sites, other_sites = ico.sites[:64], ico.sites[64:160], ico.sites[160:224]

    srv = ctx.server(4096, 1 << 20)
    r_grid = srv.submit(g_mesh, grid_sites, sides_m, seed=31, distance_grid=True, **kw)
    r_exact = srv.submit(g_mesh, exact_sites, sides_m, seed=32, **kw)
    r_k2 = srv.submit(g_mesh, other_sites, sides_m, seed=33, distance_grid=True, grid_kappa=2.0, **kw)
    mixed = _drain(srv)
    srv.close()

    exact = ctx.estimate(g_mesh, exact_sites, sides_m, seed=32, **kw)
    grid = ctx.estimate(g_mesh, grid_sites, sides_m, seed=31, distance_grid=True, **kw)
    k2 = ctx.estimate(g_mesh, other_sites, sides_m, seed=33, distance_grid=True, grid_kappa=2.0, **kw)
    for rid, ref in ((r_exact, exact), (r_grid, grid), (r_k2, k2)):
        for k in KEYS:
            np.testing.assert_array_equal(getattr(mixed[rid], k), getattr(ref, k), err_msg=k)
    # The threshold changes the paths (so it reached the kernel).
    k4 = ctx.estimate(g_mesh, other_sites, sides_m, seed=33, distance_grid=True, **kw)
    assert not np.array_equal(k4.steps, k2.steps)


def test_admission_waits_for_slots(env):
    ctx, ico, g_mesh, ball, g_ball = env
    srv = ctx.server(100, 1 << 20)
    kw = dict(combine=twos.CROSS, precision=32, walks=64, eps=1e-4, max_steps=8192)
    ids = [srv.submit(g_ball, ball.sites[:60], W.sides(ball, twos.side), seed=s, **kw) for s in range(4)]
    queued, active = srv.pending()
    assert queued + active == 4
    done = _drain(srv)
    assert sorted(done) == sorted(ids)
    with pytest.raises(twos.AdmissionError):
        srv.submit(g_ball, np.zeros((101, 3)), W.sides(ball, twos.side), seed=1, **kw)
    srv.close()


def test_insertion_validation_sees_the_cavity(env):
    ctx, ico, g_mesh, ball, g_ball = env
    # A site deep enough that every probe (within 2.5 r) starts inside the mesh.
    d, _, _ = ctx.closest(g_mesh, ico.sites, precision=64)
    z = ico.sites[int(np.argmax(d > 0.12))]
    r = 0.03
    g2 = ctx.union_sphere(g_mesh, z, r)
    rng = np.random.default_rng(0)
    d = rng.normal(size=(64, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    near = z + d * 1.5 * r
    kw = dict(combine=twos.SQUARE, precision=32, seed=77, walks=4096, eps=1e-4, max_steps=8192)
    prim = W.sides(ico, twos.side)[:1]
    base = ctx.estimate(g_mesh, near, prim, **kw)
    cav = ctx.estimate(g2, near, prim, **kw)
    # u = x + 0.3 y^2 > 0 on the mesh; a zero-data cavity can only lower u
    # (maximum principle), and with shared streams every probe sees it.
    change = base.u - cav.u
    assert np.all(change > 0)
    se = np.sqrt(np.maximum(base.same - base.u ** 2, 0) / base.walks)
    assert np.mean(change / se) > 5
    ctx.release_geometry(g2)
