"""The benchmarked configuration on an MI355X: the launch shape and distance
grid that the committed tuning cache picks for the 81,920-triangle
SH-icosphere workload (binary32 CROSS, eps 1e-4), read from that file.

- Step batching is a schedule choice only: every step_batch variant at the
  tuned shape gives bitwise the same per-site results as no batching.
- The tuned grid (binary16 corner records at the tuned resolution, threshold
  down to 2 eps from the surface) estimates what the exact walk estimates.
- The whole benchmark configuration (adaptive 256 -> 16,384 walks per side to
  1.96 SE <= 2 %|TD|, hipGraph rounds) agrees with the CPU oracle's exact
  binary64 walk.
- The grid's lower bound never exceeds the exact distance at that
  resolution, also within a few eps of the surface.
The query the grid stands in for is the reference's exact closest-point query
(benchmarks/compiled_bunny/src/wih_twos.cpp:339-359) inside walk()
(:635-664)."""

import numpy as np
import pytest

import oracle as O
from tuned_shape import SHAPE_KEYS, tuned_choice

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402

KEYS = ("u", "v", "score", "score_se", "td", "same", "walks", "steps", "caps", "status")


@pytest.fixture(scope="module")
def env():
    choice = tuned_choice()
    w = W.sh_icosphere()
    ctx = twos.Context()
    mesh = ctx.mesh(w.vertices, w.faces)
    geom = ctx.geometry(twos.MESH, mesh=mesh)
    print("tuned choice", choice, flush=True)
    ctx.set_distance_grid(mesh, choice["grid_cells"], choice["grid_corners"])
    shape = {k: choice[k] for k in SHAPE_KEYS if k in choice}
    yield ctx, w, mesh, geom, choice, shape
    ctx.close()


def test_tuned_choice_is_the_benchmarked_one():
    """The regime these tests cover: binary16 corner records, a fine grid and a
    threshold below a quarter cell (so 2 eps sets it), step batching on."""
    c = tuned_choice()
    assert c["grid_corners"] == 2 and c["grid_cells"] >= 512 and c["grid_kappa"] <= 0.25
    assert c["step_batch"] != 0


def _kw(w, shape, **over):
    kw = dict(combine=twos.CROSS, precision=32, eps=w.eps, max_steps=w.max_steps, distance_grid=True, **shape)
    kw.update(over)
    return kw


@pytest.mark.parametrize("variant", [dict(step_batch=1), dict(step_batch=24), dict(step_batch=32),
                                     dict(step_batch=40), dict(step_batch=1, waves_per_simd=0),
                                     dict(step_batch=1, block=64), dict(step_batch=1, block=256),
                                     dict(step_batch=32, block=256, waves_per_simd=0), dict(refill_batch=1),
                                     dict(refill_batch=4), dict(refill_batch=16), dict(refill_batch=64),
                                     dict(step_batch=1, refill_batch=32, block=64)])
def test_step_batching_never_changes_results(env, variant):
    ctx, w, _, geom, _, shape = env
    sites = w.sites[::64]  # 256 sites spread over the workload
    sides = W.sides(w, twos.side)
    ladder = dict(seed=3, walks=256, walks_max=1024, ci_rel=0.02)
    ref = ctx.estimate(geom, sites, sides, **_kw(w, shape, step_batch=0), **ladder)
    got = ctx.estimate(geom, sites, sides, **_kw(w, shape, **variant), **ladder)
    assert got.stats["walk_launches"] >= 1 and np.any(ref.walks > 256)  # the ladder ran more than one rung
    for k in KEYS:
        np.testing.assert_array_equal(getattr(got, k), getattr(ref, k), err_msg=f"{variant} {k}")


def test_tuned_graph_rounds_equal_eager(env):
    ctx, w, _, geom, _, shape = env
    sites = w.sites[1::64]
    sides = W.sides(w, twos.side)
    kw = _kw(w, shape, seed=5, walks=256, walks_max=2048, ci_rel=0.02)
    eager = ctx.estimate(geom, sites, sides, **kw)
    for _ in range(3):
        graph = ctx.estimate(geom, sites, sides, use_graph=True, **kw)
        for k in KEYS:
            np.testing.assert_array_equal(getattr(graph, k), getattr(eager, k), err_msg=k)
        ctx.estimate(geom, w.sites[:16], sides, **kw)  # eager work between replays


def test_tuned_grid_against_exact_walk(env):
    """At the tuned resolution and threshold, 128 sites x 4096 walks per side:
    the grid walk and the exact walk (independent seeds) agree within
    combined standard errors, with no systematic shift."""
    ctx, w, _, geom, _, shape = env
    sites = w.sites[::128]
    sides = W.sides(w, twos.side)
    exact = ctx.estimate(geom, sites, sides, **_kw(w, shape, distance_grid=False), seed=1, walks=4096)
    grid = ctx.estimate(geom, sites, sides, **_kw(w, shape), seed=2, walks=4096)
    z = (grid.td - exact.td) / np.hypot(grid.score_se, exact.score_se)
    print("grid vs exact: max |z|", np.max(np.abs(z)), "mean z", np.mean(z), flush=True)
    assert len(z) == 128 and np.all(np.isfinite(z))
    assert np.sum(np.abs(z) > 3) <= 2 and np.max(np.abs(z)) < 5
    assert abs(np.mean(z)) < 0.35  # sd of the mean of 128 unit normals is 0.09
    # The grid answers most queries (fewer traversals), so paths differ.
    assert not np.array_equal(grid.steps, exact.steps)


def test_benchmark_configuration_against_oracle(env):
    """24 sites in the benchmark's exact configuration (tuned shape and grid,
    adaptive ladder to 1.96 SE <= 2 %|TD|, hipGraph rounds) against the
    oracle's exact binary64 walk at 8192 walks per side."""
    ctx, w, _, geom, _, shape = env
    sites = w.sites[5::683][:24]
    sides = W.sides(w, twos.side)
    gpu = ctx.estimate(geom, sites, sides, **_kw(w, shape), seed=2026, walks=w.walks, walks_max=w.walks_max,
                       ci_rel=w.ci_rel, use_graph=True)
    assert np.all(gpu.status == 0) and np.all(1.96 * gpu.score_se <= w.ci_rel * np.abs(gpu.score))
    cpu = O.estimate(O.Geometry(O.MESH, mesh=O.Mesh(w.vertices, w.faces)), sites, W.sides(w, O.side),
                     combine=O.CROSS, rng=O.RNG_PHILOX, seed=77, walks=8192, eps=w.eps, max_steps=w.max_steps)
    z = (gpu.td - cpu["td"]) / np.hypot(gpu.score_se, cpu["score_se"])
    print("benchmark config vs oracle: z", np.round(z, 2).tolist(), flush=True)
    assert len(z) == 24 and np.all(np.isfinite(z))
    assert np.sum(np.abs(z) > 3) <= 1 and np.max(np.abs(z)) < 5


def _near_surface(ctx, geom, w, n, seed):
    """Points at distances 1e-5 .. 2e-3 from the surface, inside and outside
    (from random points moved along the ray to their closest surface point)."""
    rng = np.random.default_rng(seed)
    lo, hi = w.vertices.min(0), w.vertices.max(0)
    p = rng.uniform(lo, hi, (n, 3))
    d, q, _ = ctx.closest(geom, p, precision=64)
    u = (p - q) / d[:, None]
    step = np.exp(rng.uniform(np.log(1e-5), np.log(2e-3), n))
    return q + u * step[:, None]


@pytest.mark.parametrize("precision", [32, 64])
def test_tuned_grid_bound_is_conservative(env, precision):
    """The tuned grid's lower bound never exceeds the binary64 exact distance:
    20,000 points in and around the mesh and 20,000 within 2e-3 of its
    surface (where the walk's threshold 2 eps sits); and far from the
    surface it is within 1.5 cells of the distance."""
    ctx, w, _, geom, choice, _ = env
    rng = np.random.default_rng(11)
    lo, hi = w.vertices.min(0), w.vertices.max(0)
    pts = np.vstack([rng.uniform(lo - 0.05, hi + 0.05, (20000, 3)), _near_surface(ctx, geom, w, 20000, 12)])
    bound = ctx.distance_bound(geom, pts, precision=precision)
    exact, _, _ = ctx.closest(geom, pts, precision=64)
    assert np.all(bound <= exact), (np.max(bound - exact), precision)
    far = exact > 0.1
    assert np.median((exact - bound)[far]) < 1.5 * (hi - lo).max() / choice["grid_cells"]
