"""Certified minimum-benefit sampling (twos_certify; cert_walk and
run_minimum_benefit, benchmarks/compiled_bunny/src/minimum_benefit.hpp:
122-188) on an MI355X, through the C-ABI, against the pilot's retained
--mode validate samples (benchmarks/minimum_benefit/pilot-floor-policy/).

- every retained attempt (six on the watertight bunny with the affine
  source, six on the housing with its chips; up to five cavities): the first
  2,048 samples' step counts and capped / clipped flags exactly and values
  to 1e-10 (bitwise on most: the device library's cos / sin may differ from
  glibc's by an ulp, and over up to 128 steps that moves a few values by
  about 1e-12 relative);
- all 65,536 samples of two attempts: each stage's path_caps and clipped
  counts exactly, its mean to 1e-11 (summed in long double, as the
  reference does);
- the certified clearance of every finalist bitwise;
- a later block of samples (sample_begin) equals the same samples of one
  large call, so stages can be run one at a time.
"""

import numpy as np
import pytest

import cert_fixtures as CF

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def setup(ctx, golden):
    meshes = {k: ctx.mesh(*CF.mesh(golden, k)) for k in ("bunny", "housing")}
    return meshes, CF.load(golden)


def _certify(ctx, meshes, a, walks, sample_begin=0):
    g = ctx.geometry(twos.MESH, mesh=meshes[a.kind], cavities=a.holes if len(a.holes) else None,
                     cavity_radius=a.radius)
    try:
        return ctx.certify(g, a.point, seed=a.seed, walks=walks, max_steps=a.max_steps, eps=a.epsilon,
                           radius_cap=a.radius_cap, value_cap=a.value_cap,
                           source=twos.CERT_AFFINE if a.kind == "bunny" else twos.CERT_HOUSING, slope=a.slope,
                           sample_begin=sample_begin)
    finally:
        ctx.release_geometry(g)


def test_first_samples_of_every_attempt(ctx, setup):
    meshes, attempts = setup
    for a in attempts:
        print("certify", a.source, len(a.holes), "cavities", flush=True)
        value, steps, flags = _certify(ctx, meshes, a, len(a.value))
        np.testing.assert_array_equal(steps[0], a.steps, err_msg=a.source)
        np.testing.assert_array_equal(flags[0] & twos.CERT_CAPPED, a.capped, err_msg=a.source)
        np.testing.assert_array_equal((flags[0] & twos.CERT_CLIPPED) >> 1, a.clipped, err_msg=a.source)
        assert not np.any(flags & twos.CERT_CONTAINMENT)
        np.testing.assert_allclose(value[0], a.value, rtol=1e-10, atol=0, err_msg=a.source)
        assert np.mean(value[0] == a.value) > 0.5, a.source  # a systematic error would leave none bitwise


@pytest.mark.parametrize("index", [0, 9])
def test_all_stages_of_an_attempt(ctx, setup, index):
    meshes, attempts = setup
    a = attempts[index]
    print("certify all samples", a.source, flush=True)
    value, steps, flags = _certify(ctx, meshes, a, a.walks)
    for st in a.stages:
        n = st["n"]
        assert int(np.sum(flags[0, :n] & twos.CERT_CAPPED)) == st["path_caps"]
        assert int(np.sum((flags[0, :n] & twos.CERT_CLIPPED) >> 1)) == st["clipped"]
        mean = float(np.sum(value[0, :n].astype(np.longdouble)) / n)
        assert mean == pytest.approx(st["mean"], rel=1e-11, abs=0)
        assert float(np.mean(steps[0, :n])) == pytest.approx(st["mean_steps"], rel=1e-12)


def test_certified_clearance(ctx, setup):
    meshes, attempts = setup
    for a in attempts:
        g = ctx.geometry(twos.MESH, mesh=meshes[a.kind], cavities=a.holes if len(a.holes) else None,
                         cavity_radius=a.radius)
        try:
            got = ctx.cert_distance(g, a.point)[0]
        finally:
            ctx.release_geometry(g)
        assert got == a.stages[0]["center_clearance_lower"], a.source


def test_sample_blocks_compose(ctx, setup):
    meshes, attempts = setup
    a = attempts[3]
    whole = _certify(ctx, meshes, a, 1536)
    tail = _certify(ctx, meshes, a, 512, sample_begin=1024)
    for x, y in zip(whole, tail):
        np.testing.assert_array_equal(x[:, 1024:], y)
