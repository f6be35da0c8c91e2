"""The autotuner's search on a synthetic cost surface (no GPU): coordinate
descent reaches the surface's minimum over the valid shapes, never picks an
invalid variant, and caches by key."""

import numpy as np
import pytest

from twos import autotune


class FakeContext:
    device = 0

    def __init__(self, cost):
        self.cost = cost
        self.calls = 0
        self.cells = 128
        self.cell_history = []
        self.batches = []
        self.batch_factor = {}
        self.refills = []
        self.refill_factor = {}

    def set_distance_grid(self, mesh, cells, corners=False):
        self.cells = cells
        self.corners = corners
        self.cell_history.append(cells)
        return 0.0

    def estimate(self, geom, pts, sides, *, block, waves_per_simd, lds_stack, grid_kappa, step_batch=0,
                 refill_batch=0, **kw):
        self.calls += 1
        self.batches.append((step_batch, waves_per_simd, lds_stack))
        self.refills.append(refill_batch)
        ms = self.cost(block, waves_per_simd, lds_stack, grid_kappa) * (1.0 + 0.1 * abs(self.cells - 192) / 64)
        ms *= self.batch_factor.get(step_batch, 1.0) * self.refill_factor.get(refill_batch, 1.0)

        class R:
            stats = {"walk_kernel_ms": ms * 1000.0, "paths": 1000}
        return R()


def _surface(block, waves, lds, kappa):
    # Separable bowl with its minimum at (128, 5, 16, 3).
    return (1.0 + 0.1 * abs(np.log2(block / 128)) + 0.05 * abs(waves - 5) * (waves != 0) + 0.2 * (waves == 0)
            + 0.03 * abs(lds - 16) / 8 + 0.03 * abs(kappa - 3))


def test_descent_finds_the_minimum():
    autotune.clear_cache()
    ctx = FakeContext(_surface)
    t = autotune.autotune(ctx, 0, np.zeros((10, 3)), None, geom_sig="fake", combine=2, distance_grid=True,
                          repeats=1)
    assert t.choice == {"block": 128, "waves_per_simd": 5, "lds_stack": 16, "step_batch": 0, "refill_batch": 8,
                        "grid_kappa": 3.0}
    assert t.speedup > 1.0
    # Step batching exists only for the 24-entry stack with 0 or 5 waves.
    assert all(b == 0 or (lds == 24 and w in (0, 5)) for b, w, lds in ctx.batches)


def test_step_batch_is_chosen_when_faster():
    autotune.clear_cache()
    ctx = FakeContext(lambda b, w, l, k: 1.0 + 0.02 * (l != 24) + 0.05 * (w != 5))
    ctx.batch_factor = {1: 0.85, 32: 0.9}
    t = autotune.autotune(ctx, 0, np.zeros((10, 3)), None, geom_sig="fake6", combine=2, repeats=1)
    assert t.choice["step_batch"] == 1 and t.choice["lds_stack"] == 24 and t.choice["waves_per_simd"] == 5


def test_gains_below_one_percent_keep_the_default():
    autotune.clear_cache()
    t = autotune.autotune(FakeContext(lambda b, w, l, k: 1.0 - 0.005 * (b == 64)), 0, np.zeros((4, 3)), None,
                          geom_sig="fake5", combine=2, repeats=1)
    assert t.choice == {k: autotune.DEFAULTS[k] for k in ("block", "waves_per_simd", "lds_stack", "step_batch",
                                                          "refill_batch", "grid_kappa")}


def test_invalid_variants_are_never_run():
    autotune.clear_cache()
    seen = []

    def cost(block, waves, lds, kappa):
        seen.append((waves, lds))
        return 1.0 + 0.1 * (lds != 32) + 0.1 * (waves != 6)

    autotune.autotune(FakeContext(cost), 0, np.zeros((4, 3)), None, geom_sig="fake2", combine=2, repeats=1)
    # lds 16/32 exist only for waves 0 and 5.
    assert all(lds == 24 or waves in (0, 5) for waves, lds in seen)


def test_refill_batch_is_chosen_when_faster():
    autotune.clear_cache()
    ctx = FakeContext(lambda b, w, l, k: 1.0)
    ctx.refill_factor = {16: 0.9, 32: 0.95}
    t = autotune.autotune(ctx, 0, np.zeros((10, 3)), None, geom_sig="fake7", combine=2, repeats=1)
    assert t.choice["refill_batch"] == 16
    assert set(ctx.refills) == set(autotune.AXES["refill_batch"])


def test_binary64_tunes_block_only():
    autotune.clear_cache()
    seen = []

    def cost(block, waves, lds, kappa):
        seen.append((block, waves, lds))
        return 1.0 / block

    t = autotune.autotune(FakeContext(cost), 0, np.zeros((4, 3)), None, geom_sig="fake3", combine=2,
                          precision=64, repeats=1)
    assert t.choice["block"] == 256
    assert {(w, l) for _, w, l in seen} == {(0, 24)}


def test_memory_cache_skips_trials():
    autotune.clear_cache()
    ctx = FakeContext(_surface)
    autotune.autotune(ctx, 0, np.zeros((4, 3)), None, geom_sig="fake4", combine=2, repeats=1)
    n = ctx.calls
    autotune.autotune(ctx, 0, np.zeros((4, 3)), None, geom_sig="fake4", combine=2, repeats=1)
    assert ctx.calls == n


def test_grid_resolution_is_tuned_and_left_applied():
    autotune.clear_cache()
    ctx = FakeContext(lambda b, w, l, k: 1.0)
    t = autotune.autotune(ctx, 0, np.zeros((4, 3)), None, geom_sig="fake6", combine=2, distance_grid=True,
                          mesh=1, repeats=1)
    assert t.choice["grid_cells"] == 192 and ctx.cells == 192
    ctx.set_distance_grid(1, 96)
    autotune.autotune(ctx, 0, np.zeros((4, 3)), None, geom_sig="fake6", combine=2, distance_grid=True,
                      mesh=1, repeats=1)
    assert ctx.cells == 192  # a cached choice is applied again
