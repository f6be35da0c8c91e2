"""The CPU oracle against closed forms, published vectors and brute force.

Each check is chosen so that a plausible slip (a dropped term, a transposed
operand, a wrong sign or index) fails it, rather than restating the formula.
"""

import itertools
import math

import numpy as np
import pytest

import oracle as O


def test_philox_known_answers():
    # Random123 known-answer vectors for Philox4x32-10 (Salmon et al., SC'11).
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0])) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                        0x24126EA1]


def test_splitmix64_and_mt19937_64_known_answers():
    # SplitMix64 from state 0 yields 0xE220A8397B1DCDAF first.
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    # The C++ standard fixes the 10000th output of mt19937_64 with seed 5489.
    assert int(O.mt19937_64(5489, 10000)[-1]) == 9981545732273789042
    # uniform_real_distribution<double>(0,1) of a 64-bit engine: x / 2^64.
    x = O.mt19937_64(12345, 8)
    u = O.mt_uniforms(12345, 8)
    np.testing.assert_array_equal(u, np.array([float(int(v)) / 2.0**64 for v in x]))


def test_philox_key_distinct_streams():
    keys = {tuple(O.philox_key(7, s, side)) for s in range(512) for side in (0, 1)}
    assert len(keys) == 1024


def _brute_cross(a, b):
    n = len(a)
    total = sum(a[i] * b[j] for i in range(n) for j in range(n) if i != j)
    return total / (n * (n - 1))


@pytest.mark.parametrize("square", [True, False])
def test_combine_cross_index_and_jackknife_brute_force(square):
    rng = np.random.default_rng(3)
    for n in (4, 5, 9, 17):
        a = rng.normal(0.3, 1.0, n)
        b = a.copy() if square else rng.normal(-0.7, 2.0, n) + 0.5 * a
        out = O.combine(a, None if square else b, square=square)
        assert out[2] == pytest.approx(_brute_cross(a, b), rel=1e-12, abs=1e-15)
        # Explicit leave-one-out jackknife.
        loo = [_brute_cross(np.delete(a, i), np.delete(b, i)) for i in range(n)]
        se = math.sqrt((n - 1) / n * sum((x - np.mean(loo)) ** 2 for x in loo))
        assert out[3] == pytest.approx(se, rel=1e-10)
        assert out[0] == pytest.approx(a.mean()) and out[1] == pytest.approx(b.mean())
        assert out[4] == pytest.approx(np.mean(a * b))


def test_combine_powersums_equals_leave_one_out():
    rng = np.random.default_rng(11)
    for n, square in itertools.product((4, 7, 64, 4096), (True, False)):
        a = rng.exponential(1.0, n)
        b = a if square else rng.normal(0.2, 0.5, n)
        s = [a.sum(), b.sum(), (a * b).sum(), (a * a).sum(), (b * b).sum(), (a * a * b).sum(),
             (a * b * b).sum(), (a * a * b * b).sum()]
        ref = O.combine(a, b, square=False)
        got = O.combine_powersums(s, n)
        np.testing.assert_allclose(got[:3], ref[:3], rtol=1e-12)
        assert got[3] == pytest.approx(ref[3], rel=1e-7)


def test_combine_below_four_walks_has_no_se():
    out = O.combine([1.0, 2.0, 3.0], square=True)
    assert math.isnan(out[3])


def _tri_samples(a, b, c, k=60):
    pts = []
    for i in range(k + 1):
        for j in range(k + 1 - i):
            u, v = i / k, j / k
            pts.append(a + u * (b - a) + v * (c - a))
    return np.array(pts)


def test_closest_point_triangle_optimality():
    rng = np.random.default_rng(5)
    cp = np.zeros(3)
    for _ in range(300):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        O.lib().og_closest_point_triangle(*[O._ptr(np.ascontiguousarray(x), O._dp)
                                            for x in (p, a, b, c, cp)])
        # q lies in the triangle: barycentric coordinates in [0,1].
        m = np.column_stack([b - a, c - a])
        uv, *_ = np.linalg.lstsq(m, cp - a, rcond=None)
        assert uv.min() >= -1e-9 and uv.sum() <= 1 + 1e-9
        # Variational inequality (p - q).(x - q) <= 0 at the vertices.
        for x in (a, b, c):
            assert np.dot(p - cp, x - cp) <= 1e-9
        # No sampled point of the triangle is closer.
        d = np.linalg.norm(_tri_samples(a, b, c) - p, axis=1).min()
        assert np.linalg.norm(p - cp) <= d + 1e-12


def icosphere(level: int):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts), np.array(f, dtype=np.uint32)


@pytest.fixture(scope="module")
def sphere_mesh():
    v, f = icosphere(3)
    return O.Mesh(v, f), v, f


def test_bvh_equals_brute_force(sphere_mesh):
    mesh, v, f = sphere_mesh
    rng = np.random.default_rng(2)
    pts = rng.normal(size=(400, 3)) * rng.uniform(0.05, 1.5, (400, 1))
    for p in pts:
        d1, q1, t1, *_ = mesh.closest(p)
        d2, q2, t2, *_ = mesh.closest(p, brute=True)
        assert d1 == d2 and t1 == t2 and np.array_equal(q1, q2)


def test_mesh_distance_close_to_sphere_distance(sphere_mesh):
    mesh, v, f = sphere_mesh
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(100, 3))
    pts *= (rng.uniform(0, 0.8, 100) / np.linalg.norm(pts, axis=1))[:, None]
    for p in pts:
        d, *_ = mesh.closest(p)
        # Level-3 icosphere chords sag by < 1e-2 below the unit sphere.
        assert 1 - np.linalg.norm(p) - 0.01 <= d <= 1 - np.linalg.norm(p) + 1e-12


def test_ray_parity_inside(sphere_mesh):
    mesh, v, f = sphere_mesh
    rng = np.random.default_rng(6)
    dirs = rng.normal(size=(300, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    r = rng.uniform(0, 1.6, 300)
    pts = dirs * r[:, None]
    got = mesh.inside(pts)
    keep = np.abs(r - 1) > 0.02
    assert np.array_equal(got[keep], (r[keep] < 1).astype(np.int32))


def test_analytic_box_spheres_cylinders():
    spheres = np.array([[0.3, 0.3, 0.3, 0.1], [0.7, 0.6, 0.5, 0.05]])
    cyl = np.array([[2, 0.5, 0.8, 0.05]])  # along z through (x, y) = (0.5, 0.8)
    geo = O.Geometry(O.BOX, box_lo=(0, 0, 0), box_hi=(1, 1, 1), spheres=spheres, cylinders=cyl)
    g = geo.struct()
    rng = np.random.default_rng(8)
    q = np.zeros(3)
    cls = O.C.c_int32()
    for p in rng.uniform(0, 1, (500, 3)):
        pc = np.ascontiguousarray(p)
        d = O.lib().og_closest_boundary(O.C.byref(g), O._ptr(pc, O._dp), O._ptr(q, O._dp),
                                        O.C.byref(cls))
        cand = [p[0], 1 - p[0], p[1], 1 - p[1], p[2], 1 - p[2]]
        cand += [np.linalg.norm(p - s[:3]) - s[3] for s in spheres]
        cand += [math.hypot(p[0] - 0.5, p[1] - 0.8) - 0.05]
        assert d == pytest.approx(min(cand), abs=1e-15)
        # The returned point is on the boundary and at distance |d|.
        assert np.linalg.norm(q - p) == pytest.approx(abs(d), abs=1e-12)
        k = int(np.argmin(cand))
        assert cls.value == (k if k < 6 else 6)


def _ball_geometry():
    return O.Geometry(O.BALL, ball=(0.0, 0.0, 0.0, 1.0))


def test_unit_ball_green_mass():
    """-Lap u = 1 in B(0,1), u = 0 on the sphere: u(0) = 1/6, and the
    concentric-cavity compliance TD is -u(0)^2 = -1/36
    (experiments/verify_minimum_benefit.py:60-66,
    benchmarks/exact_cavity_audit/audit.py:11-25)."""
    one = O.side(source=(1.0, 0, 0, 0, 0, 0, 0))
    r = O.estimate(_ball_geometry(), [[0.0, 0.0, 0.0]], [one], combine=O.SQUARE,
                   rng=O.RNG_PHILOX, seed=9, walks=20000, eps=1e-4, max_steps=4096)
    # Each path telescopes (|x_k|^2 is a martingale-like sum of r^2), so the
    # spread is tiny; eps bias is O(eps).
    assert abs(r["u"][0] - 1 / 6) < 5e-4
    assert abs(r["score"][0] - 1 / 36) < 3 * r["score_se"][0] + 2e-4


@pytest.mark.parametrize("rng", [O.RNG_PHILOX, O.RNG_MT_SITE])
def test_green_point_moments(rng):
    """A walk from the centre of the unit ball with eps = 0.99 takes exactly one
    step (r = 1, then it lands on the sphere), so its value is f(a) / 6 at one
    Green point a = t omega with t ~ Beta(2,2): E|a|^2 = 3/10, E|a|^4 = 1/7,
    E a = 0, E (a.c)^2 = |c|^2 / 10.  For the bump f = w (1 - |a - c|^2 / R^2)^2
    with R large enough to cover the ball, E f = w (1 - 2 (3/10 + |c|^2) / R^2
    + (1/7 + |c|^2 + |c|^4) / R^4).  A uniform t, a wrong radius scale or a
    biased direction each miss this by far more than the sampling error."""
    c = np.array([0.3, -0.2, 0.1])
    w, R = 1.5, 2.0
    bump = O.side(terms=[(O.TERM_BUMP, *c, R, 0.0, 0.0, w)])
    n = 200000
    r = O.estimate(_ball_geometry(), [[0.0, 0.0, 0.0]], [bump], combine=O.SQUARE, rng=rng, seed=4,
                   walks=n, eps=0.99, max_steps=8, dump=True)
    assert r["steps"][0] == n  # one step per walk
    vals = r["a"][0]
    c2 = float(c @ c)
    expect = w * (1 - 2 * (0.3 + c2) / R**2 + (1 / 7 + c2 + c2**2) / R**4) / 6
    se = vals.std() / math.sqrt(n)
    assert abs(vals.mean() - expect) < 4 * se, (vals.mean(), expect, se)
    # The uniform-t alternative (E|a|^2 = 1/3, E|a|^4 = 1/5, E (a.c)^2 = |c|^2 / 9)
    # is excluded.
    wrong = w * (1 - 2 * (1 / 3 + c2) / R**2 + (1 / 5 + 4 * c2 / 9 + 2 * c2 / 3 + c2**2) / R**4) / 6
    assert abs(vals.mean() - wrong) > 10 * se


def test_grad_combine_brute_force_and_jackknife():
    """grad_dot counts every ordered pair of distinct walks once: equal to the
    O(N^2) sum over i != j of a_i . b_j / (N(N-1)); its first-order SE tracks
    the leave-one-out jackknife of that U-statistic for large N."""
    rng = np.random.default_rng(5)
    for n, corr in ((7, 0.0), (40, 0.6)):
        a = rng.normal(size=(n, 3)) + [0.5, -1.0, 2.0]
        b = corr * a + rng.normal(size=(n, 3)) + [1.0, 0.3, -0.2]
        out = O.combine_grad(a, b)
        brute = sum(a[i] @ b[j] for i in range(n) for j in range(n) if i != j) / (n * (n - 1))
        assert abs(out[6] - brute) < 1e-12 * max(1.0, abs(brute))
        np.testing.assert_allclose(out[:3], a.mean(0), rtol=1e-14)
        np.testing.assert_allclose(out[3:6], b.mean(0), rtol=1e-14)
    n = 3000
    a = rng.normal(size=(n, 3)) + [0.5, -1.0, 2.0]
    b = 0.4 * a + rng.normal(size=(n, 3)) + [1.0, 0.3, -0.2]
    out = O.combine_grad(a, b)
    A, B, Cs = a.sum(0), b.sum(0), np.einsum("ij,ij->", a, b)
    loo = ((A - a) * (B - b)).sum(1) - (Cs - np.einsum("ij,ij->i", a, b))
    loo /= (n - 1) * (n - 2)
    jack = np.sqrt((n - 1) / n * np.sum((loo - loo.mean()) ** 2))
    assert abs(out[7] / jack - 1) < 0.02
    assert np.isnan(O.combine_grad(a[:3], b[:3])[7])  # no SE below four walks


def test_gradient_channel_c1_exact():
    """Configuration 1: grad g_u . grad g_v = 2xz - 2y + 0.5x and the mean
    gradients (2x, -2y, 0.5), (z, 1, x) from the first-step weights."""
    gu = O.side(poly=[0, 0, 0, 0.5, 1, -1, 0, 0, 0, 0], poly_classes=(0,))
    gv = O.side(poly=[0, 0, 1, 0, 0, 0, 0, 0, 1, 0], poly_classes=(0,))
    z = np.array([[0.1, 0.2, -0.3], [0.4, -0.2, 0.1], [0.0, 0.0, 0.0], [-0.5, 0.3, 0.2], [0.2, 0.5, 0.4]])
    r = O.estimate(_ball_geometry(), z, [gu, gv], combine=O.CROSS, rng=O.RNG_PHILOX, seed=1, walks=20000,
                   eps=1e-3, max_steps=4096, gradient=True, threads=8)
    x, y, w = z.T
    err = (r["grad_dot"] - (2 * x * w - 2 * y + 0.5 * x)) / r["grad_dot_se"]
    assert np.max(np.abs(err)) < 4, err
    np.testing.assert_allclose(r["grad_u"], np.c_[2 * x, -2 * y, 0.5 + 0 * x], atol=0.08)
    np.testing.assert_allclose(r["grad_v"], np.c_[w, 1 + 0 * x, x], atol=0.08)
    # Without the flag the channel is NaN.
    r0 = O.estimate(_ball_geometry(), z[:1], [gu, gv], combine=O.CROSS, seed=1, walks=64, eps=1e-3)
    assert np.isnan(r0["grad_dot"][0]) and np.all(np.isnan(r0["grad_u"][0]))


def test_gradient_channel_affine_source():
    """-Lap u = f, u = 0 on the unit sphere, f = f0 + g.x: u = f0 (1-|x|^2)/6 +
    (g.x)(1-|x|^2)/10, so grad u = -f0 x/3 + g (1-|x|^2)/10 - (g.x) x/5; the
    first-step sample needs the deterministic (r0^2/10) g term to get this."""
    side = O.reference_affine_side(0.65)   # f = 1 + 0.65 (y + 0.5): f0 = 1.325, g = (0, 0.65, 0)
    z = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1], [-0.1, 0.4, 0.2]])
    r = O.estimate(_ball_geometry(), z, [side], combine=O.SQUARE, rng=O.RNG_PHILOX, seed=2, walks=40000,
                   eps=1e-4, max_steps=4096, gradient=True, threads=8)
    f0, g = 1.325, np.array([0.0, 0.65, 0.0])
    q = 1 - np.sum(z * z, 1)
    exact = -f0 * z / 3 + np.outer(q, g) / 10 - np.outer(z @ g, np.ones(3)) * z / 5
    # Sampling error per component is about 0.0015 at these sites (40000 walks).
    np.testing.assert_allclose(r["grad_u"], exact, atol=0.008)
    # At the centre the first sphere reaches the boundary, so the estimate is
    # exactly (r0^2/10) g; with r0^2/6 it would be 0.108.
    np.testing.assert_allclose(r["grad_u"][0], [0.0, 0.065, 0.0], atol=1e-12)


def test_unit_ball_harmonic_td_c1():
    """Configuration 1: g_u = x^2 - y^2 + 0.5 z, g_v = x z + y are harmonic, so
    u = g_u, v = g_v and the Dirichlet-cavity TD is -g_u g_v
    (paper/sections/03_problem.tex:153 with h = 0)."""
    gu = O.side(poly=[0, 0, 0, 0.5, 1, -1, 0, 0, 0, 0], poly_classes=(0,))
    gv = O.side(poly=[0, 0, 1, 0, 0, 0, 0, 0, 1, 0], poly_classes=(0,))
    rng = np.random.default_rng(0)
    d = rng.normal(size=(16, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    z = d * (0.8 * rng.uniform(0, 1, 16) ** (1 / 3))[:, None]
    r = O.estimate(_ball_geometry(), z, [gu, gv], combine=O.CROSS, rng=O.RNG_PHILOX, seed=1,
                   walks=4096, eps=1e-3, max_steps=4096)
    g_u = z[:, 0] ** 2 - z[:, 1] ** 2 + 0.5 * z[:, 2]
    g_v = z[:, 0] * z[:, 2] + z[:, 1]
    err = np.abs(r["td"] - (-g_u * g_v)) / r["score_se"]
    assert np.mean(err < 3.5) >= 0.9 and np.max(err) < 5


def test_paired_and_cross_differ_only_in_streams():
    gu = O.side(poly=[0, 1, 0, 0, 0, 0, 0, 0, 0, 0], poly_classes=(0,))
    z = [[0.2, -0.1, 0.3]]
    p = O.estimate(_ball_geometry(), z, [gu, gu], combine=O.PAIRED, rng=O.RNG_PHILOX, seed=5,
                   walks=256, eps=1e-3, dump=True)
    # Paired: one path carries both accumulators, so a == b per walk.
    np.testing.assert_array_equal(p["a"], p["b"])
    c = O.estimate(_ball_geometry(), z, [gu, gu], combine=O.CROSS, rng=O.RNG_PHILOX, seed=5,
                   walks=256, eps=1e-3, dump=True)
    np.testing.assert_array_equal(c["a"], p["a"])      # side 0 uses the side-0 stream
    assert not np.array_equal(c["b"], p["b"])           # side 1 has its own stream


def test_adaptive_ladder_meets_target_or_caps():
    gu = O.side(poly=[0.3, 1, 0, 0, 0, 0, 0, 0, 0, 0], poly_classes=(0,))
    gv = O.side(poly=[0.5, 0, 0, 1, 0, 0, 0, 0, 0, 0], poly_classes=(0,))
    z = np.array([[0.1, 0.0, 0.2], [0.0, 0.1, -0.5], [0.5, 0.5, 0.1]])
    r = O.estimate(_ball_geometry(), z, [gu, gv], combine=O.CROSS, rng=O.RNG_PHILOX, seed=3,
                   walks=64, walks_max=4096, ci_rel=0.05, eps=1e-3)
    for i in range(3):
        met = 1.96 * r["score_se"][i] <= 0.05 * abs(r["score"][i])
        assert (r["status"][i] == 0) == met
        assert r["walks"][i] in (64, 128, 256, 512, 1024, 2048, 4096)
        if r["status"][i] == 1:
            assert r["walks"][i] == 4096


def test_adaptive_ladder_stops_at_the_first_rung_that_meets_the_target():
    """Minimality: each path depends only on (seed, site, side, walk), so a
    fixed-budget run of W walks reproduces the ladder's state at rung W.  A
    site that stopped at rung W met the target there and did not meet it at
    W / 2; a site that ran out of budget met it at no rung.  A ladder that
    never stops early, or one that stops a rung late, fails this."""
    gu = O.side(poly=[0, 0, 0, 0.5, 1, -1, 0, 0, 0, 0], poly_classes=(0,))
    gv = O.side(poly=[0, 0, 1, 0, 0, 0, 0, 0, 1, 0], poly_classes=(0,))
    rng = np.random.default_rng(0)
    d = rng.normal(size=(24, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    z = d * (0.8 * rng.uniform(0, 1, 24) ** (1 / 3))[:, None]
    kw = dict(combine=O.CROSS, rng=O.RNG_PHILOX, seed=12, eps=1e-3, max_steps=8192)
    rungs = [64, 128, 256, 512, 1024, 2048]
    r = O.estimate(_ball_geometry(), z, [gu, gv], walks=64, walks_max=2048, ci_rel=0.3, **kw)
    fixed = {w: O.estimate(_ball_geometry(), z, [gu, gv], walks=w, **kw) for w in rungs}
    met = {w: 1.96 * fixed[w]["score_se"] <= 0.3 * np.abs(fixed[w]["score"]) for w in rungs}
    stops = set()
    for i in range(len(z)):
        w = int(r["walks"][i])
        stops.add(w)
        # The ladder's result at its last rung is the fixed-budget result there.
        assert r["score"][i] == fixed[w]["score"][i] and r["steps"][i] == fixed[w]["steps"][i]
        if r["status"][i] == 0:
            assert met[w][i], i
            assert all(not met[v][i] for v in rungs if v < w), i
        else:
            assert w == 2048 and not any(met[v][i] for v in rungs), i
    assert len(stops) >= 3 and 2048 in stops and min(stops) < 2048  # the sites spread over the rungs
