"""The greedy insertion design loop on an MI355X.

- GpuStudy on the reference's paired streams (binary64) replays the pilot
  design (benchmarks/submission_study/pilot/pipeline/branch/twos-1611): the
  six accepted centres equal design.json's, with the recorded shortlists
  consistent with the GPU's scout scores (see test_design_cpu.py).
- The production path (CROSS on Philox streams, binary32 walks) runs the same
  loop, and its choice validates: the accepted cavity lowers J at
  sensor-distributed probes by about its predicted benefit u p x 4 pi r.
"""

import json

import numpy as np
import pytest

from twos import design as D

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


@pytest.fixture(scope="module")
def env(golden):
    d = json.loads((golden / "study_branch_design.json").read_text())
    m = np.load(golden / "study_branch_mesh.npz")
    ctx = twos.Context()
    mesh = ctx.mesh(m["vertices"], m["faces"])
    yield ctx, mesh, d
    ctx.close()


def test_pilot_design_replay_on_gpu(env):
    ctx, mesh, d = env
    r0 = d["rounds"][0]
    est = D.GpuStudy(ctx, mesh, d["config"], radius=d["radius"], eps=r0["epsilon"], max_steps=r0["max_steps"])
    cfg = D.GreedyConfig(rounds=len(d["centers"]), radius=d["radius"], seed=d["seed"],
                         scout_walks=r0["scout"]["walks"], rescore_walks=r0["rescore"]["walks"])

    def recorded(k, scores):
        rec = np.asarray(d["rounds"][k]["shortlist"])
        cutoff = np.sort(scores)[-len(rec)]
        assert np.all(scores[rec] >= cutoff) and set(np.nonzero(scores > cutoff)[0]) <= set(rec), k
        return rec

    out = D.greedy(est, np.array(d["pool"]), cfg, shortlist_fn=recorded)
    for k, (got, want) in enumerate(zip(out.rounds, d["rounds"])):
        print("round", k, len(got.feasible), flush=True)
        assert len(got.feasible) == want["scout"]["sites"]
        ref = np.asarray(want["rescore"]["score"])
        scale = np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref)) + 1e-300)
        assert np.all(np.abs(got.rescore_score - ref) <= 1e-6 * scale), k
    np.testing.assert_array_equal(out.centers, np.array(d["centers"]))


def test_production_design_and_validation(env):
    """Current source configuration (with the sensor self-heating bump), CROSS
    on Philox streams: two rounds on the pilot's pool; the first choice is
    validated with coupled fresh-seed estimates of J at 256 probes."""
    ctx, mesh, d = env
    config = "0.57 0.37 0 0.1\n-0.66 0 0 0.15 35\n0.57 -0.38 0 0.13 45\n0.57 0.37 0 0.085 4.0\n"
    est = D.GpuStudy(ctx, mesh, config, radius=0.065, eps=2e-4, max_steps=4096, streams="philox")
    cfg = D.GreedyConfig(rounds=2, radius=0.065, seed=7, scout_walks=512, rescore_walks=8192, shortlist=8,
                         select="lcb")
    out = D.greedy(est, np.array(d["pool"]), cfg)
    assert len(out.centers) == 2 and np.linalg.norm(out.centers[0] - out.centers[1]) > 2 * 0.065 + 0.025
    first = out.rounds[0]
    best = int(np.argmax(first.rescore_score - 1.96 * first.rescore_se))
    predicted, predicted_se = first.rescore_score[best], first.rescore_se[best]
    assert predicted > 3 * predicted_se  # a clearly beneficial site
    probes, w = D.ball_probes([0.57, 0.37, 0.0], 0.1, 256, seed=11)
    v = D.validate_insertion(est.u, probes, w, [], out.centers[0], 0.065, walks=8192, seed=99)
    print("predicted", predicted, "+-", predicted_se, "validated", v.per_capacity, "+-", v.per_capacity_se,
          flush=True)
    # A finite ball's benefit per capacity is close to the point prediction
    # u p (first order in r); both are noisy, so only the sign and the order
    # of magnitude are checked.
    assert v.delta_j > 0 and v.per_capacity > 3 * v.per_capacity_se
    assert 0.2 < v.per_capacity / predicted < 5


def test_c3_design_loop_with_cavities_and_distance_grid():
    """Configuration 5 in small on the 81,920-triangle SH icosphere, at the
    launch shape and distance grid the committed tuning cache picks for it:
    two rounds over 4096 candidates (scout 512 walks per side, the top 8
    rescored with 4096 fresh walks, LCB acceptance, radius-0.02 cavities), so
    round 2 walks the mesh minus the first cavity with the grid on.  Each
    round's rescoring agrees with the oracle's exact binary64 walk on that
    geometry within combined SEs, and a binary64 GPU rerun on the same
    Philox streams differs from the binary32 rescoring by under 0.5 SE
    (run_study.py:77-106; internal_cooling/run_reference_benchmark.py:196-245)."""
    import oracle as O
    from twos import workloads as W
    from tuned_shape import tuned_choice, tuned_shape

    choice, shape = tuned_choice(), tuned_shape()
    w = W.sh_icosphere(n_sites=4096)
    with twos.Context() as ctx:
        mesh = ctx.mesh(w.vertices, w.faces)
        ctx.set_distance_grid(mesh, choice["grid_cells"], choice["grid_corners"])
        sides = W.sides(w, twos.side)
        est = D.GpuDesign(ctx, mesh, sides, radius=0.02, eps=w.eps, max_steps=w.max_steps, shape=shape)
        cfg = D.GreedyConfig(rounds=2, radius=0.02, clearance=0.025, scout_walks=512, shortlist=8,
                             rescore_walks=4096, seed=2026, select="lcb")
        out = D.greedy(est, w.sites, cfg)
        assert len(out.centers) == 2 and len(out.rounds[1].feasible) < len(out.rounds[0].feasible)
        est64 = D.GpuDesign(ctx, mesh, sides, radius=0.02, eps=w.eps, max_steps=w.max_steps, precision=64,
                            shape=shape)
        omesh = O.Mesh(w.vertices, w.faces)
        for k, r in enumerate(out.rounds):
            pts = w.sites[r.feasible][r.shortlist]
            holes = out.centers[:k]
            geo = O.Geometry(O.MESH, mesh=omesh, cavities=holes if k else None, cavity_radius=0.02)
            cpu = O.estimate(geo, pts, W.sides(w, O.side), combine=O.CROSS, rng=O.RNG_PHILOX, seed=991 + k,
                             walks=8192, eps=w.eps, max_steps=w.max_steps)
            z = (r.rescore_score - cpu["score"]) / np.hypot(r.rescore_se, cpu["score_se"])
            s64, se64 = est64(pts, holes, cfg.rescore_walks, cfg.seed + cfg.seed_step * k + cfg.rescore_offset)
            dev = np.abs(r.rescore_score - np.asarray(s64)) / np.asarray(se64)
            print("round", k, "z vs oracle", np.round(z, 2).tolist(), "binary64 rerun", float(dev.max()), flush=True)
            assert np.all(np.isfinite(z)) and np.sum(np.abs(z) > 3) <= 1 and np.max(np.abs(z)) < 5
            assert np.max(dev) < 0.5
