"""twos_estimate_device on an MI355X: sites and results in device memory
(torch tensors' data_ptr()), ordered on the caller's stream.  The results
equal twos_estimate's on the same host sites bit for bit, the estimate waits
for the work that produces the sites on that stream, and later work on the
stream sees the results."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import workloads as W  # noqa: E402

KEYS = ("u", "v", "score", "score_se", "td", "same", "walks", "steps", "caps", "status")


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    ctx = twos.Context()
    ico = W.sh_icosphere(n_sites=600)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    yield torch, ctx, ico, ctx.geometry(twos.MESH, mesh=mesh)
    ctx.close()


def _device_estimate(torch, ctx, geom, sites, sides, **kw):
    stream = torch.cuda.Stream()
    n = len(sites)
    src = torch.tensor(sites, dtype=torch.float64).pin_memory()
    d_sites = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    d_out = torch.zeros((n * twos.RESULT_BYTES,), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        if hasattr(torch.cuda, "_sleep"):
            torch.cuda._sleep(20_000_000)  # the sites arrive late on this stream
        d_sites.copy_(src, non_blocking=True)
        stats = ctx.estimate_device(geom, d_sites.data_ptr(), n, d_out.data_ptr(), sides,
                                    stream=stream.cuda_stream, **kw)
        back = d_out.to("cpu", non_blocking=False)  # queued on the same stream
    stream.synchronize()
    return twos.results_from_bytes(back.numpy().tobytes(), n), stats


@pytest.mark.parametrize("grid", [False, True])
def test_device_pointers_equal_host_arrays(env, grid):
    torch, ctx, ico, g = env
    sides = W.sides(ico, twos.side)
    kw = dict(combine=twos.CROSS, precision=32, seed=21, walks=64, walks_max=1024, ci_rel=0.05, eps=ico.eps,
              max_steps=ico.max_steps, distance_grid=grid)
    host = ctx.estimate(g, ico.sites, sides, **kw)
    dev, stats = _device_estimate(torch, ctx, g, ico.sites, sides, **kw)
    for k in KEYS:
        np.testing.assert_array_equal(getattr(dev, k), getattr(host, k), err_msg=k)
    assert stats["steps"] == int(host.steps.sum()) and stats["paths"] == host.stats["paths"]


def test_device_estimate_reports_step_caps(env):
    torch, ctx, ico, g = env
    sides = W.sides(ico, twos.side)
    with pytest.raises(twos.StepCapError):
        _device_estimate(torch, ctx, g, ico.sites[:8], sides, combine=twos.CROSS, walks=16, eps=ico.eps,
                         max_steps=2)
