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


def test_graph_rounds_survive_geometry_table_growth(env):
    """A server capturing its rounds as hipGraphs keeps equal results while
    more than 16 and 32 derived geometries are created (the device geometry
    table is reallocated and every engine's captured rounds are replaced)."""
    ctx, _, _, ball, g_ball = env
    srv = ctx.server(1024, 1 << 18)
    kw = dict(combine=twos.CROSS, precision=32, walks=256, eps=1e-3, max_steps=8192)
    sides = W.sides(ball, twos.side)
    made = []
    try:
        for i in range(40):
            c = [0.5 * np.cos(i), 0.5 * np.sin(i), 0.0]
            g = ctx.union_sphere(g_ball, c, 0.05)
            made.append(g)
            rid = srv.submit(g, ball.sites[:32], sides, seed=40 + i, use_graph=True, **kw)
            got = _drain(srv)[rid]
            ref = ctx.estimate(g, ball.sites[:32], sides, seed=40 + i, **kw)
            for k in KEYS:
                np.testing.assert_array_equal(getattr(got, k), getattr(ref, k), err_msg=f"{i} {k}")
    finally:
        srv.close()
        for g in made:
            ctx.release_geometry(g)


def test_geometry_in_use_cannot_be_released(env):
    ctx, _, _, ball, g_ball = env
    g = ctx.union_sphere(g_ball, [0.3, 0.0, 0.0], 0.05)
    srv = ctx.server(1024, 1 << 18)
    kw = dict(combine=twos.CROSS, precision=32, walks=64, walks_max=1024, ci_rel=0.01, eps=1e-3, max_steps=8192)
    srv.submit(g, ball.sites[:16], W.sides(ball, twos.side), seed=3, **kw)
    with pytest.raises(twos.InvalidArgument):
        ctx.release_geometry(g)  # queued
    srv.step()
    if srv.pending() != (0, 0):
        with pytest.raises(twos.InvalidArgument):
            ctx.release_geometry(g)  # active
    _drain(srv)
    ctx.release_geometry(g)
    srv.close()


def test_failed_request_is_reported_and_frees_its_slots(env):
    """A request whose paths hit max_steps ends with ERR_STEP_CAP (the
    reference's serve loop answers DONE 1 and exits, study.cpp:373-374); the
    requests around it complete with their standalone results, and its slots
    serve the next request."""
    ctx, ico, g_mesh, ball, g_ball = env
    srv = ctx.server(256, 1 << 20)
    kw = dict(combine=twos.CROSS, precision=32, walks=128, eps=1e-4, max_steps=8192)
    sides_m, sides_b = W.sides(ico, twos.side), W.sides(ball, twos.side)
    ra = srv.submit(g_mesh, ico.sites[:64], sides_m, seed=1, distance_grid=True, **kw)
    rb = srv.submit(g_mesh, ico.sites[64:128], sides_m, seed=2, **{**kw, "max_steps": 3})
    rc = srv.submit(g_ball, ball.sites[:64], sides_b, seed=3, **kw)
    done = _drain(srv)
    assert done[rb].stats["status"] == twos.StepCapError.code and np.all(done[rb].caps > 0)
    for rid, g, pts, sides, seed, extra in ((ra, g_mesh, ico.sites[:64], sides_m, 1, {"distance_grid": True}),
                                            (rc, g_ball, ball.sites[:64], sides_b, 3, {})):
        assert done[rid].stats["status"] == 0
        ref = ctx.estimate(g, pts, sides, seed=seed, **kw, **extra)
        for k in KEYS:
            np.testing.assert_array_equal(getattr(done[rid], k), getattr(ref, k), err_msg=k)
    assert srv.pending() == (0, 0)
    rd = srv.submit(g_ball, ball.sites[:64].repeat(4, axis=0), sides_b, seed=4, **kw)  # every slot
    assert _drain(srv)[rd].stats["status"] == 0
    srv.close()


def test_admission_by_memory_counts_resident_geometry(golden):
    """Admission charges each request's slots against the context's HBM budget
    beside the resident geometry: with room for 1.5 requests, the second
    waits for the first although slots are free; once a mesh with its
    distance grid takes the room, a request that can never fit is reported
    with ERR_ADMISSION while the server goes on."""
    kw = dict(combine=twos.CROSS, precision=32, walks=64, eps=1e-3, max_steps=8192)
    ball = W.unit_ball(n_sites=64)
    sides = W.sides(ball, twos.side)
    with twos.Context() as probe:
        probe.geometry(twos.BALL)
        srv = probe.server(2048, 1 << 16)
        m = probe.memory()
        srv.close()
    slot = m["slot_bytes"]
    budget = m["geometry_bytes"] + m["engine_bytes"] + 1500 * slot
    with twos.Context(hbm_budget=budget) as ctx:
        g = ctx.geometry(twos.BALL)
        srv = ctx.server(2048, 1 << 16)
        assert ctx.memory()["engine_bytes"] == m["engine_bytes"]
        sites = np.tile(ball.sites, (16, 1))[:1000]
        ra = srv.submit(g, sites, sides, seed=1, **kw)
        rb = srv.submit(g, sites, sides, seed=2, **kw)
        assert srv.pending() == (2, 0)
        first = srv.step()
        assert set(first) == {ra} and srv.pending() == (1, 0)  # 2000 sites' slots are free, 1500 sites' bytes
        second = _drain(srv)
        assert set(second) == {rb}
        with twos.Context() as free:  # standalone references, outside the tight budget
            gf = free.geometry(twos.BALL)
            for rid, res, seed in ((ra, first[ra], 1), (rb, second[rb], 2)):
                ref = free.estimate(gf, sites, sides, seed=seed, **kw)
                for k in KEYS:
                    np.testing.assert_array_equal(getattr(res, k), getattr(ref, k), err_msg=k)
        v = np.load(golden / "bunny_1000.npz")
        mesh = ctx.mesh(v["vertices"], v["faces"])  # LBVH + 128-cell grid: megabytes
        assert ctx.memory()["geometry_bytes"] > m["geometry_bytes"] + 1500 * slot
        rc = srv.submit(g, sites, sides, seed=3, **kw)
        out = srv.step()
        assert out[rc].stats["status"] == twos.AdmissionError.code and srv.pending() == (0, 0)
        ctx.release_mesh(mesh)
        rd = srv.submit(g, sites, sides, seed=4, **kw)
        assert _drain(srv)[rd].stats["status"] == 0
        srv.close()


def test_stream_run_equals_standalone_estimates():
    """Configuration 4 in small: a Poisson stream of 40 requests over the
    three resident geometries (unit ball, cooling block, SH icosphere),
    validations included, replayed through serving.run.  Every request
    finishes, each result equals its standalone estimate bit for bit, each
    validation's derived geometry is released, and a validation's objective
    change has the sign the maximum principle gives."""
    from twos import serving

    with twos.Context() as ctx:
        stream = serving.Stream(ctx, twos, n_requests=40, rate=50.0, seed=5, max_sites=256, validation_frac=0.3)
        srv = ctx.server(max_sites=8192, max_walks_per_round=1 << 22)
        out = serving.run(stream, srv)
        srv.close()
        print({k: out[k] for k in ("rounds", "p50_ms", "p99_ms", "validations", "failed")}, flush=True)
        assert out["requests"] == 40 and out["failed"] == 0
        assert {r.geom_name for r in stream.requests if r.kind == "td"} == set(stream.geoms)
        vals = [r for r in stream.requests if r.kind == "validation"]
        assert len(vals) >= 3
        for req in stream.requests:
            assert np.isfinite(req.finish) and len(req.results) == 1
            (_, got), = req.results
            g = stream.derived_geometry(req) if req.kind == "validation" else stream.geoms[req.geom_name][1]
            geom, sites, sides, kw = stream.request_args(req, g)
            ref = ctx.estimate(geom, sites, sides, **kw)
            for k in KEYS:
                np.testing.assert_array_equal(getattr(got, k), getattr(ref, k), err_msg=f"{req.rid} {req.kind} {k}")
            if req.kind == "validation":
                ctx.release_geometry(g)
                with pytest.raises(twos.NotFoundError):
                    ctx.release_geometry(req.extra["derived"])  # released by serving.run
        # The cooling block's and the icosphere's primal data are >= 0 (u = 1 on
        # holes, 0 on faces; x + 0.3 y^2 on the mesh in the unit box) and the
        # cavity holds u = 0, so by the maximum principle cutting it can only
        # lower J (the unit ball's harmonic data changes sign).
        signed = [serving.validation_change(r) for r in vals if r.geom_name != "unit_ball"]
        for change, se in signed:
            assert change > -3 * se, (change, se)
