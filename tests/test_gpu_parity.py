"""HIP library parity on an MI355X, through the C-ABI.

Checkers: the CPU oracle (itself pinned bit-exactly to the reference's retained
traces and fields) and the reference's own fixtures.
"""

import json

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

twos = pytest.importorskip("twos")


@pytest.fixture(scope="module")
def ctx():
    c = twos.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def bunny(golden):
    meta = json.loads((golden / "bunny_meta.json").read_text())
    sites = np.load(golden / "bunny_sites.npy")
    meshes = {k: np.load(golden / f"bunny_{k}.npz") for k in (1000, 69674)}
    return meta, sites, meshes


@pytest.fixture(scope="module")
def gpu_bunny(ctx, bunny):
    meta, sites, meshes = bunny
    out = {}
    for k, m in meshes.items():
        mesh = ctx.mesh(m["vertices"], m["faces"])
        out[k] = (mesh, ctx.geometry(twos.MESH, mesh=mesh), ctx.geometry(twos.MESH, mesh=mesh, mesh_radius_f32=True))
    return out


def _box_points(v, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = v.min(0), v.max(0)
    pad = 0.1 * (hi - lo)
    return rng.uniform(lo - pad, hi + pad, (n, 3))


@pytest.mark.parametrize("faces", [1000, 69674])
def test_lbvh_closest_binary64_equals_brute_force(ctx, bunny, gpu_bunny, faces):
    _, _, meshes = bunny
    v, f = meshes[faces]["vertices"], meshes[faces]["faces"]
    om = O.Mesh(v, f)
    pts = _box_points(v, 1500, faces)
    d, q, c = ctx.closest(gpu_bunny[faces][1], pts, precision=64)
    for i in range(0, len(pts), 1 if faces == 1000 else 7):
        dr, qr, *_ = om.closest(pts[i], brute=faces == 1000)
        assert d[i] == dr, i
        assert np.array_equal(q[i], qr), i
    assert np.all(c == 0)


def test_lbvh_closest_binary32_close(ctx, bunny, gpu_bunny):
    _, _, meshes = bunny
    v, f = meshes[69674]["vertices"], meshes[69674]["faces"]
    om = O.Mesh(v, f)
    pts = _box_points(v, 800, 3)
    d, q, _ = ctx.closest(gpu_bunny[69674][1], pts, precision=32)
    ref = np.array([om.closest(p)[0] for p in pts])
    assert np.max(np.abs(d - ref)) < 2e-6


def test_mesh_info_depth_within_stack(ctx, gpu_bunny):
    mesh = gpu_bunny[69674][0]
    assert 17 <= mesh.depth < 64
    assert mesh.n_faces == 69674


def test_walk_traces_replay_on_gpu(ctx, bunny, gpu_bunny, golden):
    """GPU binary64 walks on the reference's own mt19937_64 streams reproduce
    benchmarks/compiled_bunny/output/walk-traces.csv."""
    meta, sites, _ = bunny
    traces = np.load(golden / "bunny_walk_traces.npy")
    g = gpu_bunny[69674][2]
    side = twos.reference_affine_side()
    worst = 0.0
    for path in range(12):
        rows = traces[traces[:, 0] == path]
        site = path * 384 // 12
        got = ctx.trace(g, sites[site], side, rng=twos.RNG_MT_KEY,
                        seed=meta["traces"]["seed"] + meta["traces"]["key_offset"], site=site, eps=5e-4,
                        max_steps=512)
        assert len(got) == len(rows), path
        worst = max(worst, float(np.abs(got - rows[:, 2:6]).max()))
    # Bitwise but for libm: glibc and the device library may round cos/sin differently by an ulp.
    assert worst <= 1e-13


@pytest.mark.parametrize("name", ["faces_1000", "twos_32", "twos_512"])
def test_field_replay_on_gpu(ctx, bunny, gpu_bunny, golden, name):
    meta, sites, _ = bunny
    f = meta["fields"][name]
    ref = np.load(golden / f"field_{name}.npz")
    res = ctx.estimate(gpu_bunny[f["faces"]][2], sites, [twos.reference_affine_side(f["source_y_slope"])],
                       combine=twos.SQUARE, rng=twos.RNG_MT_SITE, precision=64, seed=f["seed"], repeat=f["repeat"],
                       walks=f["walks"], eps=f["epsilon"], max_steps=f["max_steps"])
    # Integer step counts must agree exactly (an ulp of cos/sin could in principle
    # move one termination; allow at most one site).
    assert np.sum(res.steps != ref["steps"]) <= 1
    same = res.steps == ref["steps"]
    np.testing.assert_allclose(res.u[same], ref["u"][same], rtol=1e-11)
    np.testing.assert_allclose(res.score[same], ref["score"][same], rtol=1e-10)
    np.testing.assert_allclose(res.score_se[same], ref["score_se"][same], rtol=1e-7)
    np.testing.assert_array_equal(res.caps, ref["limit_hits"])


def _c1():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(16, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    z = d * (0.8 * rng.uniform(0, 1, 16) ** (1 / 3))[:, None]
    gu = [0, 0, 0, 0.5, 1, -1, 0, 0, 0, 0]
    gv = [0, 0, 1, 0, 0, 0, 0, 0, 1, 0]
    return z, gu, gv


@pytest.mark.parametrize("schedule", ["plain", "persistent"])
def test_philox_binary64_matches_oracle(ctx, schedule):
    z, gu, gv = _c1()
    g = ctx.geometry(twos.BALL)
    res = ctx.estimate(g, z, [twos.side(poly=gu, poly_classes=(0,)), twos.side(poly=gv, poly_classes=(0,))],
                       combine=twos.CROSS, precision=64, seed=3,
                       schedule=twos.PLAIN if schedule == "plain" else twos.PERSISTENT, walks=256, eps=1e-3,
                       max_steps=4096, dump=True)
    ref = O.estimate(O.Geometry(O.BALL), z, [O.side(poly=gu, poly_classes=(0,)), O.side(poly=gv, poly_classes=(0,))],
                     combine=O.CROSS, rng=O.RNG_PHILOX, seed=3, walks=256, eps=1e-3, max_steps=4096, dump=True)
    # Same streams, same binary64 operations; glibc and the device library may
    # round sin/cos differently by an ulp, and a path carries that forward, so
    # per-path samples agree to 1e-9 and step counts exactly.
    np.testing.assert_allclose(res.a, ref["a"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(res.b, ref["b"], rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(res.walk_steps, ref["walk_steps"])
    np.testing.assert_array_equal(res.steps, ref["steps"])
    np.testing.assert_allclose(res.score, ref["score"], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(res.score_se, ref["score_se"], rtol=1e-7)


def test_combine_matches_binary64_recombination(ctx, bunny, gpu_bunny):
    """The GPU's power-sum combine reproduces a CPU binary64 recombination of
    the same per-walk samples to 1e-12 (score) and the O(N) jackknife."""
    _, sites, _ = bunny
    gu = twos.side(poly=[0, 1, 0, 0, 0, 0.3, 0, 0, 0, 0], poly_classes=(0,))
    gv = twos.side(poly=[0, 0, 0, 1, 0, 0, 0, 0, 0, 0], poly_classes=(0,))
    res = ctx.estimate(gpu_bunny[69674][1], sites[:48], [gu, gv], combine=twos.CROSS, precision=32, seed=11,
                       walks=512, eps=1e-4, max_steps=8192, dump=True)
    for i in range(48):
        ref = O.combine(res.a[i], res.b[i])
        assert res.score[i] == pytest.approx(ref[2], rel=1e-12, abs=1e-18)
        assert res.u[i] == pytest.approx(ref[0], rel=1e-13)
        assert res.score_se[i] == pytest.approx(ref[3], rel=1e-8)


def test_plain_and_persistent_bitwise_equal(ctx, bunny, gpu_bunny):
    _, sites, _ = bunny
    gu = twos.side(poly=[0, 1, 0, 0, 0, 0.3, 0, 0, 0, 0], poly_classes=(0,))
    gv = twos.side(poly=[0, 0, 0, 1, 0, 0, 0, 0, 0, 0], poly_classes=(0,))
    kw = dict(combine=twos.CROSS, precision=32, seed=5, walks=64, walks_max=1024, ci_rel=0.05, eps=1e-4,
              max_steps=8192)
    a = ctx.estimate(gpu_bunny[69674][1], sites[:40], [gu, gv], schedule=twos.PLAIN, **kw)
    b = ctx.estimate(gpu_bunny[69674][1], sites[:40], [gu, gv], schedule=twos.PERSISTENT, **kw)
    for k in ("u", "v", "score", "score_se", "walks", "steps", "caps", "status"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def test_binary32_against_binary64_rerun(ctx):
    z, gu, gv = _c1()
    g = ctx.geometry(twos.BALL)
    sides = [twos.side(poly=gu, poly_classes=(0,)), twos.side(poly=gv, poly_classes=(0,))]
    r32 = ctx.estimate(g, z, sides, combine=twos.CROSS, precision=32, seed=9, walks=2048, eps=1e-3, max_steps=4096)
    r64 = ctx.estimate(g, z, sides, combine=twos.CROSS, precision=64, seed=9, walks=2048, eps=1e-3, max_steps=4096)
    # Same Philox streams: binary32 paths shadow the binary64 ones, so the
    # estimates differ far less than their standard errors.
    se = np.hypot(r32.score_se, r64.score_se)
    assert np.max(np.abs(r32.score - r64.score) / se) < 0.5


@pytest.mark.parametrize("faces", [1000, 69674])
def test_inside_matches_oracle(ctx, bunny, gpu_bunny, faces):
    """Ray parity (EmbreeMesh::inside, wih_twos.cpp:361-394) is a boolean: the
    GPU's binary64 Moller-Trumbore hits along the fixed direction, with the
    same tnear, tfar and 2e-6 advance, equal the oracle's at every point."""
    _, _, meshes = bunny
    v, f = meshes[faces]["vertices"], meshes[faces]["faces"]
    pts = _box_points(v, 4000, 9)
    got = ctx.inside(gpu_bunny[faces][0], pts)
    ref = O.Mesh(v, f).inside(pts)
    bad = np.nonzero(got != (ref == 1))[0]
    assert len(bad) == 0, (bad[:10].tolist(), pts[bad[:10]].tolist())
    assert 0.1 < np.mean(got) < 0.9  # both classes occur


def test_graph_replays_equal_eager(ctx, bunny, gpu_bunny):
    _, sites, _ = bunny
    gu = twos.side(poly=[0, 1, 0, 0, 0, 0.3, 0, 0, 0, 0], poly_classes=(0,))
    gv = twos.side(poly=[0, 0, 0, 1, 0, 0, 0, 0, 0, 0], poly_classes=(0,))
    kw = dict(combine=twos.CROSS, precision=32, seed=21, walks=64, walks_max=2048, ci_rel=0.03, eps=1e-4,
              max_steps=8192)
    eager = ctx.estimate(gpu_bunny[69674][1], sites[:64], [gu, gv], **kw)
    for _ in range(3):
        graph = ctx.estimate(gpu_bunny[69674][1], sites[:64], [gu, gv], use_graph=True, **kw)
        for k in ("score", "score_se", "walks", "steps"):
            np.testing.assert_array_equal(getattr(graph, k), getattr(eager, k), err_msg=k)
        ctx.estimate(gpu_bunny[69674][1], sites[64:72], [gu, gv], **kw)  # eager work between replays
