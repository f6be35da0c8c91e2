"""Edge cases of the estimator's entry points on an MI355X, through the C-ABI.

- empty site and point lists: every entry point returns at once with empty
  results (the reference's loops simply do nothing, wih_twos.cpp:1828-1836);
- the smallest walk counts: N = 2 and 3 give the cross-index score but no
  jackknife SE (score_se NaN below four walks, wih_twos.cpp:698-699), N >= 4
  both, equal to the oracle's binary64 results;
- sites already inside the eps-shell: every path ends at its first query
  with the payoff at the closest boundary point and no step (walk(),
  wih_twos.cpp:644-647);
- more sites than one call's slots and ragged site lists: one large call
  equals the same sites split into calls with their global site offsets, so
  chunking never changes a stream.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402

GU = [0, 0, 0, 0.5, 1, -1, 0, 0, 0, 0]  # x^2 - y^2 + 0.5 z
GV = [0, 0, 1, 0, 0, 0, 0, 0, 1, 0]     # x z + y


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


def _ball_sites(n, seed=0, rmax=0.8):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d * (rmax * rng.uniform(0, 1, n) ** (1 / 3))[:, None]


def test_empty_inputs(ctx, golden):
    none = np.zeros((0, 3))
    ball = ctx.geometry(twos.BALL)
    sides = [twos.side(poly=GU, poly_classes=(0,)), twos.side(poly=GV, poly_classes=(0,))]
    res = ctx.estimate(ball, none, sides, combine=twos.CROSS, walks=16, walks_max=64, eps=1e-3)
    assert len(res.u) == 0 and len(res.score) == 0
    m = np.load(golden / "bunny_1000.npz")
    mesh = ctx.mesh(m["vertices"], m["faces"])
    c = m["vertices"].mean(0)
    g = ctx.geometry(twos.MESH, mesh=mesh, cavities=[c], cavity_radius=0.01)
    try:
        for kw in (dict(), dict(rng=twos.RNG_PHILOX, precision=32, distance_grid=True)):
            v, cnt = ctx.differential(g, none, twos.side(poly=GU, poly_classes=(0,)), walks=8, **kw)
            assert v.shape == (0, 1, 8) and cnt.shape == (0, 9)
        for kw in (dict(), dict(philox=True, grid_cells=64)):
            val, st, fl = ctx.certify(g, none, walks=8, **kw)
            assert val.shape == st.shape == fl.shape == (0, 8)
        assert len(ctx.cert_distance(g, none)) == 0
        assert len(ctx.cert_grid_lower(mesh, 64, none)) == 0
        assert len(ctx.inside(mesh, none)) == 0
        assert len(ctx.closest(g, none)[0]) == 0
    finally:
        ctx.release_geometry(g)


@pytest.mark.parametrize("walks", [2, 3, 4, 5])
@pytest.mark.parametrize("combine", ["SQUARE", "CROSS"])
def test_small_walk_counts_match_oracle(ctx, walks, combine):
    sites = _ball_sites(6, seed=3)
    g = ctx.geometry(twos.BALL)
    kw = dict(seed=9, walks=walks, eps=1e-3, max_steps=4096)
    gpu = ctx.estimate(g, sites, [twos.side(poly=GU, poly_classes=(0,)), twos.side(poly=GV, poly_classes=(0,))],
                       combine=getattr(twos, combine), precision=64, **kw)
    ref = O.estimate(O.Geometry(O.BALL), sites, [O.side(poly=GU, poly_classes=(0,)), O.side(poly=GV, poly_classes=(0,))],
                     combine=getattr(O, combine), rng=O.RNG_PHILOX, **kw)
    np.testing.assert_array_equal(gpu.steps, ref["steps"])
    np.testing.assert_allclose(gpu.score, ref["score"], rtol=1e-9, atol=1e-13)
    assert np.all(gpu.walks == walks)
    if walks < 4:
        assert np.all(np.isnan(gpu.score_se)) and np.all(np.isnan(ref["score_se"]))
    else:
        np.testing.assert_allclose(gpu.score_se, ref["score_se"], rtol=1e-7, atol=1e-13)


@pytest.mark.parametrize("precision", [32, 64])
def test_sites_in_the_eps_shell_take_no_step(ctx, precision):
    """Sites at distance eps/2 from the unit sphere: every path ends at its
    first query, so no step is taken and each sample is g at the closest
    sphere point; the estimate is that value."""
    eps = 1e-3
    d = _ball_sites(8, seed=5)
    d /= np.linalg.norm(d, axis=1)[:, None]
    sites = d * (1.0 - 0.5 * eps)
    g = ctx.geometry(twos.BALL)
    res = ctx.estimate(g, sites, [twos.side(poly=GU, poly_classes=(0,))], combine=twos.SQUARE, precision=precision,
                       walks=32, eps=eps, max_steps=64)
    assert np.all(res.steps == 0)
    y = sites / np.linalg.norm(sites, axis=1)[:, None]
    gu = y[:, 0] ** 2 - y[:, 1] ** 2 + 0.5 * y[:, 2]
    tol = 1e-12 if precision == 64 else 2e-6
    np.testing.assert_allclose(res.u, gu, rtol=0, atol=tol)
    ref = O.estimate(O.Geometry(O.BALL), sites, [O.side(poly=GU, poly_classes=(0,))], combine=O.SQUARE,
                     rng=O.RNG_PHILOX, seed=0, walks=32, eps=eps, max_steps=64)
    assert np.all(ref["steps"] == 0)
    np.testing.assert_allclose(res.u, ref["u"], rtol=0, atol=tol)


def test_large_and_ragged_site_lists_keep_their_streams(ctx):
    """9,000 sites in one call equal the same sites in calls of 5,000, 1 and
    3,999 sites at their global offsets (ladder on, so the sites stop on
    different rungs), bit for bit."""
    ico = W.sh_icosphere(n_sites=9000)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    g = ctx.geometry(twos.MESH, mesh=mesh)
    sides = W.sides(ico, twos.side)
    kw = dict(combine=twos.CROSS, seed=13, walks=32, walks_max=256, ci_rel=0.05, eps=1e-4, max_steps=8192,
              distance_grid=True)
    full = ctx.estimate(g, ico.sites, sides, **kw)
    parts = [(0, 5000), (5000, 5001), (5001, 9000)]
    for a, b in parts:
        part = ctx.estimate(g, ico.sites[a:b], sides, site_offset=a, **kw)
        for k in ("u", "v", "score", "score_se", "walks", "steps", "caps", "status"):
            np.testing.assert_array_equal(getattr(part, k), getattr(full, k)[a:b], err_msg=f"{k} {a}:{b}")
    assert len(np.unique(full.walks)) > 1  # the ladder stopped sites on different rungs
