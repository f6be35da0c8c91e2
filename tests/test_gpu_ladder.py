"""The adaptive walk ladder on an MI355X against the CPU oracle: binary64 walks
on the same Philox streams, doubling each site's walks per side until
ci_z * SE <= ci_rel * |score| or walks_max (the budget ladder of
benchmarks/matched_accuracy/adaptive.py:29-57, doubling as the 256 -> 16,384
rule of the triangle-mesh workload).  Every site stops at the same rung with
the same step count on the GPU as in the oracle, whose own ladder is checked
for minimality in test_oracle_math.py."""

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


def _compare(gpu, cpu, rungs):
    np.testing.assert_array_equal(gpu.walks, cpu["walks"])
    np.testing.assert_array_equal(gpu.status, cpu["status"])
    np.testing.assert_array_equal(gpu.steps, cpu["steps"])
    np.testing.assert_array_equal(gpu.caps, cpu["caps"])
    np.testing.assert_allclose(gpu.score, cpu["score"], rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(gpu.score_se, cpu["score_se"], rtol=1e-7)
    # The comparison is only as strong as the spread of stopping rungs.
    stopped = set(gpu.walks.tolist())
    print("stopping rungs", sorted(stopped), flush=True)
    assert len(stopped) >= 3 and stopped <= set(rungs)
    assert np.any(gpu.status == 0) and np.all(gpu.walks[gpu.status == 1] == rungs[-1])


@pytest.mark.parametrize("schedule", [twos.PERSISTENT, twos.PLAIN])
def test_unit_ball_ladder_equals_oracle(ctx, schedule):
    w = W.unit_ball(n_sites=48)
    g = ctx.geometry(twos.BALL)
    # A loose target spreads the 48 sites over every rung (the oracle stops
    # 1, 5, 5, 7, 6, 4 of them at 64 .. 2048 and caps 20 at 4096).
    kw = dict(combine=twos.CROSS, seed=12, walks=64, walks_max=4096, ci_rel=0.3, eps=w.eps, max_steps=8192)
    gpu = ctx.estimate(g, w.sites, W.sides(w, twos.side), precision=64, schedule=schedule, **kw)
    cpu = O.estimate(O.Geometry(O.BALL), w.sites, W.sides(w, O.side), rng=O.RNG_PHILOX, **kw)
    _compare(gpu, cpu, [64, 128, 256, 512, 1024, 2048, 4096])


def test_mesh_exact_ladder_equals_oracle(ctx):
    """The 81,920-triangle SH icosphere, exact closest-point radius at every
    step, binary64."""
    w = W.sh_icosphere(n_sites=512)
    mesh = ctx.mesh(w.vertices, w.faces)
    g = ctx.geometry(twos.MESH, mesh=mesh)
    sites = w.sites[::32]
    kw = dict(combine=twos.CROSS, seed=13, walks=64, walks_max=1024, ci_rel=0.05, eps=w.eps, max_steps=w.max_steps)
    gpu = ctx.estimate(g, sites, W.sides(w, twos.side), precision=64, **kw)
    cpu = O.estimate(O.Geometry(O.MESH, mesh=O.Mesh(w.vertices, w.faces)), sites, W.sides(w, O.side),
                     rng=O.RNG_PHILOX, **kw)
    _compare(gpu, cpu, [64, 128, 256, 512, 1024])
