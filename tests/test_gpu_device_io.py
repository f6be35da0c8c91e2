"""twos_estimate_device on an MI355X: sites and results in device memory,
ordered on the caller's stream (buffers and streams of the HIP runtime the
library links, tests/hip_buffers.py).  The results equal twos_estimate's on
the same host sites bit for bit; the estimate waits for the work that
produces the sites on that stream (queued behind 2 GiB of copies), and later
work on the stream sees the results."""

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
    ico = W.sh_icosphere(n_sites=600)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    busy = [HB.Buffer(1 << 30, fill=None), HB.Buffer(1 << 30, fill=None)]  # keeps a stream busy
    yield ctx, ico, ctx.geometry(twos.MESH, mesh=mesh), busy
    for b in busy:
        b.close()
    ctx.close()


def device_estimate(ctx, geom, sites, sides, busy=None, **kw):
    """The sites reach the device on a fresh stream behind `busy` copies; the
    estimate and the read-back are queued on the same stream."""
    n = len(sites)
    stream = HB.Stream()
    d_sites = HB.Buffer(24 * n)
    d_out = HB.Buffer(twos.RESULT_BYTES * n, fill=0xAB)
    try:
        if busy:
            for _ in range(2):
                busy[1].copy_from(busy[0], stream)
        d_sites.upload(np.asarray(sites, dtype=np.float64), stream)
        stats = ctx.estimate_device(geom, d_sites.ptr, n, d_out.ptr, sides, stream=stream.value, **kw)
        res = twos.results_from_bytes(d_out.download(stream), n)
    finally:
        d_sites.close()
        d_out.close()
        stream.close()
    return res, stats


@pytest.mark.parametrize("grid", [False, True])
def test_device_pointers_equal_host_arrays(env, grid):
    ctx, ico, g, busy = env
    sides = W.sides(ico, twos.side)
    kw = dict(combine=twos.CROSS, precision=32, seed=21, walks=64, walks_max=1024, ci_rel=0.05, eps=ico.eps,
              max_steps=ico.max_steps, distance_grid=grid)
    host = ctx.estimate(g, ico.sites, sides, **kw)
    dev, stats = device_estimate(ctx, g, ico.sites, sides, busy, **kw)
    for k in KEYS:
        np.testing.assert_array_equal(getattr(dev, k), getattr(host, k), err_msg=k)
    assert stats["steps"] == int(host.steps.sum()) and stats["paths"] == host.stats["paths"]


def test_device_estimate_reports_step_caps(env):
    ctx, ico, g, _ = env
    sides = W.sides(ico, twos.side)
    with pytest.raises(twos.StepCapError):
        device_estimate(ctx, g, ico.sites[:8], sides, combine=twos.CROSS, walks=16, eps=ico.eps, max_steps=2)
