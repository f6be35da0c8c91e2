"""Launch shapes on an MI355X: every block size, occupancy variant and LDS
stack depth gives bitwise the same results (paths are keyed by (seed, site,
side, walk) and folded in walk order), and the autotuner returns a valid,
no-slower shape."""

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")
from twos import autotune  # noqa: E402
from twos import workloads as W  # noqa: E402

KEYS = ("u", "v", "score", "score_se", "walks", "steps", "caps", "status")


@pytest.fixture(scope="module")
def env():
    ctx = twos.Context()
    ico = W.sh_icosphere(n_sites=512)
    mesh = ctx.mesh(ico.vertices, ico.faces)
    yield ctx, ico, ctx.geometry(twos.MESH, mesh=mesh), mesh
    ctx.close()


@pytest.mark.parametrize("grid", [False, True])
def test_launch_shape_never_changes_results(env, grid):
    ctx, ico, g, _ = env
    kw = dict(combine=twos.CROSS, precision=32, seed=3, walks=64, walks_max=256, ci_rel=0.05, eps=1e-4,
              max_steps=8192, distance_grid=grid)
    sides = W.sides(ico, twos.side)
    ref = ctx.estimate(g, ico.sites[:256], sides, **kw)
    shapes = [dict(block=64), dict(block=128), dict(block=192), dict(waves_per_simd=5), dict(waves_per_simd=6),
              dict(waves_per_simd=8), dict(lds_stack=16), dict(lds_stack=32), dict(waves_per_simd=5, lds_stack=16),
              dict(waves_per_simd=5, lds_stack=32, block=128)]
    for shape in shapes:
        print("shape", shape, flush=True)
        r = ctx.estimate(g, ico.sites[:256], sides, **kw, **shape)
        for k in KEYS:
            np.testing.assert_array_equal(getattr(r, k), getattr(ref, k), err_msg=f"{shape} {k}")


def test_autotune_choice_is_valid_and_cached(env, tmp_path):
    ctx, ico, g, mesh = env
    autotune.clear_cache()
    path = tmp_path / "tune.json"
    sides = W.sides(ico, twos.side)
    t = autotune.autotune(ctx, g, ico.sites, sides, geom_sig="sh_icosphere_512", combine=twos.CROSS,
                          precision=32, eps=1e-4, distance_grid=True, mesh=mesh, sample=512, walks=64,
                          cache_path=path)
    for axis, values in autotune.AXES.items():
        assert t.choice[axis] in values
    assert t.ms_per_path <= t.default_ms_per_path
    stored = json.loads(path.read_text())
    assert len(stored) == 1 and next(iter(stored.values()))["choice"] == t.choice
    autotune.clear_cache()
    again = autotune.autotune(ctx, g, ico.sites, sides, geom_sig="sh_icosphere_512", combine=twos.CROSS,
                              precision=32, eps=1e-4, distance_grid=True, mesh=mesh, cache_path=path)
    assert again.choice == t.choice  # read back from the file, no new trials
