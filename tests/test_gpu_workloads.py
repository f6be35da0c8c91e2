"""The three synthetic workloads on the GPU against closed forms and the
CPU oracle (statistical parity: independent Philox paths are not shared
across implementations when the radius rule differs, so agreement is in
combined standard errors)."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def c3(ctx):
    w = W.sh_icosphere(n_sites=512)
    mesh = ctx.mesh(w.vertices, w.faces)
    return w, mesh, ctx.geometry(twos.MESH, mesh=mesh)


def _z(a, b):
    return (a.td - b["td"]) / np.hypot(a.score_se, b["score_se"])


def test_unit_ball_c1_against_closed_form(ctx):
    w = W.unit_ball()
    g = ctx.geometry(twos.BALL)
    res = ctx.estimate(g, w.sites, W.sides(w, twos.side), combine=twos.CROSS, precision=64, seed=0,
                       walks=1024, eps=w.eps, max_steps=8192)
    z = (res.td - W.unit_ball_exact(w.sites)) / res.score_se
    # 64 sites: at 3 SE the expected number of misses is 0.2; allow 2.
    assert np.sum(np.abs(z) > 3) <= 2
    assert np.max(np.abs(z)) < 5


def test_c3_sites_inside_and_away_from_the_surface(ctx, c3):
    w, mesh, g = c3
    assert ctx.inside(mesh, w.sites).all()
    d, _, _ = ctx.closest(g, w.sites, precision=64)
    assert d.min() > 0.009


def test_c3_exact_walks_against_oracle(ctx, c3):
    w, mesh, g = c3
    sites = w.sites[:24]
    kw = dict(combine=twos.CROSS, seed=17, walks=2048, eps=w.eps, max_steps=w.max_steps)
    gpu = ctx.estimate(g, sites, W.sides(w, twos.side), precision=32, **kw)
    cpu = O.estimate(O.Geometry(O.MESH, mesh=O.Mesh(w.vertices, w.faces)), sites, W.sides(w, O.side),
                     rng=O.RNG_PHILOX, threads=0, **{**kw, "seed": 18})
    z = _z(gpu, cpu)
    assert np.sum(np.abs(z) > 3) <= 1 and np.max(np.abs(z)) < 5


@pytest.mark.parametrize("kappa", [0.25, 0.5, 1.0, 4.0])
def test_c3_distance_grid_against_exact(ctx, c3, kappa):
    """The conservative-radius walk estimates the same quantity as the exact
    walk (independent seeds), at every threshold the autotuner may choose."""
    w, mesh, g = c3
    sites = w.sites[:128]
    kw = dict(combine=twos.CROSS, precision=32, walks=4096, eps=w.eps, max_steps=w.max_steps)
    exact = ctx.estimate(g, sites, W.sides(w, twos.side), seed=1, **kw)
    grid = ctx.estimate(g, sites, W.sides(w, twos.side), seed=2, distance_grid=True, grid_kappa=kappa, **kw)
    z = (grid.td - exact.td) / np.hypot(grid.score_se, exact.score_se)
    assert np.sum(np.abs(z) > 3) <= 2 and np.max(np.abs(z)) < 5
    assert abs(np.mean(z)) < 0.35  # no systematic shift (128 sites: sd of the mean ~0.09)


@pytest.mark.parametrize("corners", [0, 1, 2])
def test_c3_finer_distance_grid_against_exact(ctx, c3, corners):
    """A rebuilt 256-cell grid (centre or eight-corner bound) changes the paths
    and still estimates the exact walk's quantity; the mesh is restored to its
    128-cell centre grid afterwards."""
    w, mesh, g = c3
    sites = w.sites[128:256]
    kw = dict(combine=twos.CROSS, precision=32, walks=4096, eps=w.eps, max_steps=w.max_steps)
    coarse = ctx.estimate(g, sites, W.sides(w, twos.side), seed=3, distance_grid=True, grid_kappa=2.0, **kw)
    try:
        ctx.set_distance_grid(mesh, 256, corners)
        fine = ctx.estimate(g, sites, W.sides(w, twos.side), seed=3, distance_grid=True, grid_kappa=2.0, **kw)
    finally:
        ctx.set_distance_grid(mesh, 128)
    exact = ctx.estimate(g, sites, W.sides(w, twos.side), seed=4, **kw)
    assert not np.array_equal(fine.steps, coarse.steps)
    z = (fine.td - exact.td) / np.hypot(fine.score_se, exact.score_se)
    assert np.sum(np.abs(z) > 3) <= 2 and np.max(np.abs(z)) < 5
    assert abs(np.mean(z)) < 0.35
    again = ctx.estimate(g, sites, W.sides(w, twos.side), seed=3, distance_grid=True, grid_kappa=2.0, **kw)
    np.testing.assert_array_equal(again.steps, coarse.steps)


@pytest.mark.parametrize("mode,cells", [(0, 128), (1, 256), (2, 256), (2, 96)])
@pytest.mark.parametrize("precision", [32, 64])
def test_distance_grid_bound_is_conservative(ctx, c3, mode, cells, precision):
    """At 20,000 points inside and around the mesh, the grid's lower bound
    never exceeds the exact distance (binary64 closest point), in every grid
    mode; and it is tight enough to be useful far from the boundary."""
    w, mesh, g = c3
    rng = np.random.default_rng(cells + mode)
    lo, hi = w.vertices.min(0), w.vertices.max(0)
    pts = rng.uniform(lo - 0.05, hi + 0.05, (20000, 3))
    try:
        ctx.set_distance_grid(mesh, cells, mode)
        bound = ctx.distance_bound(g, pts, precision=precision)
    finally:
        ctx.set_distance_grid(mesh, 128, 0)
    exact, _, _ = ctx.closest(g, pts, precision=64)
    assert np.all(bound <= exact), (np.max(bound - exact), mode, cells)
    far = exact > 0.1
    assert np.median((exact - bound)[far]) < 1.5 * (hi - lo).max() / cells


def test_c3_distance_grid_plain_equals_persistent(ctx, c3):
    w, mesh, g = c3
    kw = dict(combine=twos.CROSS, precision=32, seed=4, walks=64, walks_max=512, ci_rel=0.05, eps=w.eps,
              max_steps=w.max_steps, distance_grid=True)
    a = ctx.estimate(g, w.sites[:24], W.sides(w, twos.side), schedule=twos.PLAIN, **kw)
    b = ctx.estimate(g, w.sites[:24], W.sides(w, twos.side), schedule=twos.PERSISTENT, **kw)
    for k in ("score", "score_se", "walks", "steps"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def test_cooling_block_c2_against_oracle(ctx):
    w = W.cooling_block()
    g = ctx.geometry(twos.BOX, box_lo=w.box_lo, box_hi=w.box_hi, spheres=w.spheres, cylinders=w.cylinders)
    idx = np.linspace(0, len(w.sites) - 1, 32).astype(int)
    sites = w.sites[idx]
    kw = dict(combine=twos.CROSS, walks=2048, eps=w.eps, max_steps=w.max_steps)
    gpu = ctx.estimate(g, sites, W.sides(w, twos.side), precision=32, seed=5, **kw)
    geo = O.Geometry(O.BOX, box_lo=w.box_lo, box_hi=w.box_hi, spheres=w.spheres, cylinders=w.cylinders)
    cpu = O.estimate(geo, sites, W.sides(w, O.side), rng=O.RNG_PHILOX, seed=6, **kw)
    # u in [0, 1] is 1 on holes; v is the z = 1 face's harmonic measure.
    assert np.all((gpu.u >= 0) & (gpu.u <= 1)) and np.all((gpu.v >= 0) & (gpu.v <= 1))
    # A site none of whose paths reach the z = 1 face (or none a hole) has
    # every adjoint (or primal) sample 0, so TD = 0 with SE 0: where both
    # estimates are such sites there is no z-score, and their TDs must agree
    # exactly; every other site is compared in combined SEs.
    both0 = (gpu.score_se == 0) & (cpu["score_se"] == 0)
    np.testing.assert_array_equal(gpu.td[both0], cpu["td"][both0])
    np.testing.assert_array_equal(gpu.td[both0], 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        z = _z(gpu, cpu)[~both0]
    assert np.all(np.isfinite(z))
    assert len(z) >= 24, len(z)  # 32 sites, of which the oracle finds 4 with SE 0 at both seeds
    assert np.sum(np.abs(z) > 3) <= 1


def test_cooling_block_binary32_paths_shadow_binary64(ctx):
    w = W.cooling_block()
    g = ctx.geometry(twos.BOX, box_lo=w.box_lo, box_hi=w.box_hi, spheres=w.spheres, cylinders=w.cylinders)
    sites = w.sites[::97]
    kw = dict(combine=twos.CROSS, walks=1024, eps=w.eps, max_steps=w.max_steps, seed=8, dump=True)
    a = ctx.estimate(g, sites, W.sides(w, twos.side), precision=32, **kw)
    b = ctx.estimate(g, sites, W.sides(w, twos.side), precision=64, **kw)
    # Same Philox streams: most binary32 paths end on the same boundary piece
    # as their binary64 rerun.
    assert np.mean(a.a == b.a) > 0.95 and np.mean(a.b == b.b) > 0.95
