"""Failure handling forced from the host: every kernel still exits normally.

- a step cap of 1 makes every path capped: twos_estimate returns the cap
  error with the estimate attached (the reference's fatal "one or more paths
  hit --max-steps", wih_twos.cpp:1846-1848);
- a site outside a closed mesh: its paths escape, end as capped when their
  distance stops being finite or at max_steps, and the call reports the cap;
- invalid arguments are rejected before any launch.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def env():
    ctx = twos.Context()
    ico = W.sh_icosphere(n_sites=64)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    yield ctx, ico, ctx.geometry(twos.MESH, mesh=mesh)
    ctx.close()


def test_step_cap_error_carries_result(env):
    ctx, ico, g = env
    with pytest.raises(twos.StepCapError) as e:
        ctx.estimate(g, ico.sites[:8], W.sides(ico, twos.side), combine=twos.CROSS, walks=16, eps=1e-4,
                     max_steps=1)
    res = e.value.result
    assert np.all(res.caps == 2 * 16)  # both sides of every path
    res2 = ctx.estimate(g, ico.sites[:8], W.sides(ico, twos.side), combine=twos.CROSS, walks=16, eps=1e-4,
                        max_steps=1, raise_on_cap=False)
    np.testing.assert_array_equal(res2.caps, res.caps)


def test_site_outside_the_mesh_ends_capped(env):
    ctx, ico, g = env
    outside = np.array([[2.0, 2.0, 2.0]])
    with pytest.raises(twos.StepCapError) as e:
        ctx.estimate(g, outside, W.sides(ico, twos.side), combine=twos.CROSS, walks=8, eps=1e-4,
                     max_steps=4096)
    assert e.value.result.caps[0] > 0


@pytest.mark.parametrize("bad", [dict(walks=1), dict(max_steps=0), dict(eps=0.0), dict(precision=16),
                                 dict(rng=twos.RNG_MT_SITE, precision=32), dict(block=100)])
def test_invalid_arguments(env, bad):
    ctx, ico, g = env
    kw = dict(combine=twos.CROSS, walks=16, eps=1e-4, max_steps=64)
    kw.update(bad)
    with pytest.raises(twos.InvalidArgument):
        ctx.estimate(g, ico.sites[:2], W.sides(ico, twos.side), **kw)


def test_unknown_geometry(env):
    ctx, ico, g = env
    with pytest.raises(twos.NotFoundError):
        ctx.estimate(12345, ico.sites[:2], W.sides(ico, twos.side), walks=16)


def test_grid_rebuilt_before_any_geometry_exists(golden):
    """A fresh context: rebuilding a mesh's distance grid before any geometry
    is created succeeds, and a geometry created afterwards sees that grid."""
    m = np.load(golden / "bunny_1000.npz")
    with twos.Context() as ctx:
        mesh = ctx.mesh(m["vertices"], m["faces"])
        assert ctx.set_distance_grid(mesh, 64, 2) > 0.0
        g = ctx.geometry(twos.MESH, mesh=mesh)
        c = m["vertices"].mean(0)
        assert np.isfinite(ctx.distance_bound(g, [c])[0])
        ctx.mesh(m["vertices"], m["faces"])  # a later build on the same thread still works
