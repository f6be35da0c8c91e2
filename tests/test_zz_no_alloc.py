"""twos_estimate_device_static on an MI355X, through the C-ABI: the
allocation-free estimate (every buffer in the device's static scratch array)
gives the results of twos_estimate_device bit for bit, on two calls in a row,
the second with fewer sites (the array keeps the first call's contents), and
rejects walk budgets beyond its per-round capacity and graph capture with a
clear error.  Device buffers and streams come from the HIP runtime the
library links (tests/hip_buffers.py)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402

import hip_buffers as HB  # noqa: E402

KEYS = ("u", "v", "score", "score_se", "td", "same", "walks", "steps", "caps", "status")


@pytest.fixture(scope="module")
def env():
    ctx = twos.Context()
    ico = W.sh_icosphere(n_sites=5000)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    yield ctx, ico, ctx.geometry(twos.MESH, mesh=mesh)
    ctx.close()


def _run(ctx, geom, sites, sides, static, **kw):
    n = len(sites)
    stream = HB.Stream()
    d_sites = HB.Buffer(24 * n)
    d_out = HB.Buffer(twos.RESULT_BYTES * n, fill=0xAB)
    try:
        d_sites.upload(np.asarray(sites, dtype=np.float64), stream)
        ctx.estimate_device(geom, d_sites.ptr, n, d_out.ptr, sides, stream=stream.value, static=static, **kw)
        return twos.results_from_bytes(d_out.download(stream), n)
    finally:
        d_sites.close()
        d_out.close()
        stream.close()


@pytest.mark.parametrize("grid", [False, True])
def test_static_equals_device_estimate_twice(env, grid):
    ctx, ico, g = env
    sides = W.sides(ico, twos.side)
    kw = dict(combine=twos.CROSS, precision=32, walks=64, walks_max=512, ci_rel=0.05, eps=ico.eps,
              max_steps=ico.max_steps, distance_grid=grid)
    # 5000 sites: two chunks of the array's 4096 slots; then 300 sites.
    for count, seed in ((5000, 31), (300, 32)):
        sites = ico.sites[:count]
        ref = _run(ctx, g, sites, sides, False, seed=seed, **kw)
        got = _run(ctx, g, sites, sides, True, seed=seed, **kw)
        print("static", count, "sites", int(got.steps.sum()), "steps", flush=True)
        for k in KEYS:
            np.testing.assert_array_equal(getattr(got, k), getattr(ref, k), err_msg=f"{count} {k}")


def test_static_rejects_budgets_beyond_its_array(env):
    ctx, ico, g = env
    sides = W.sides(ico, twos.side)
    with pytest.raises(twos.AdmissionError, match="walks per round"):
        _run(ctx, g, ico.sites[:4], sides, True, combine=twos.CROSS, walks=64, walks_max=1 << 24, ci_rel=0.01,
             eps=ico.eps)
    with pytest.raises(twos.InvalidArgument):
        _run(ctx, g, ico.sites[:4], sides, True, combine=twos.CROSS, walks=64, eps=ico.eps, use_graph=True)
