"""Certified minimum-benefit sampling (run_minimum_benefit / cert_walk,
benchmarks/compiled_bunny/src/minimum_benefit.hpp:122-188) in the oracle,
against the pilot's retained --mode validate samples
(benchmarks/minimum_benefit/pilot-floor-policy/*/attempt-*/validation.csv).

Pins: per sample, the step count and the capped / clipped flags exactly and
the value to the last bit on >= 99% of samples (the rest within a few ulps:
glibc's cos / sin / acos may differ by an ulp from the build that wrote them,
and the walk moves by that much); the certified clearance of each finalist
bitwise; the distance equal to a brute-force minimum of the rounded-down
triangle-box distances, and never above the exact surface distance.
"""

import numpy as np
import pytest

import cert_fixtures as CF
import oracle as O


@pytest.fixture(scope="module")
def attempts(golden):
    return CF.load(golden)


@pytest.fixture(scope="module")
def meshes(golden):
    return {k: O.Mesh(*CF.mesh(golden, k)) for k in ("bunny", "housing")}


def _geometry(meshes, a):
    return O.Geometry(O.MESH, mesh=meshes[a.kind], cavities=a.holes if len(a.holes) else None,
                      cavity_radius=a.radius)


def _run(meshes, a, walks):
    return O.certify(_geometry(meshes, a), a.point, seed=a.seed, walks=walks, max_steps=a.max_steps,
                     eps=a.epsilon, radius_cap=a.radius_cap, value_cap=a.value_cap,
                     source=O.CERT_AFFINE if a.kind == "bunny" else O.CERT_HOUSING, slope=a.slope,
                     threads=O.host_threads())


@pytest.mark.parametrize("index", [0, 5, 7, 11])
def test_samples_match_reference(attempts, meshes, index):
    a = attempts[index]
    n = 768
    value, steps, flags = _run(meshes, a, n)
    np.testing.assert_array_equal(steps[0], a.steps[:n])
    np.testing.assert_array_equal(flags[0] & 1, a.capped[:n])
    np.testing.assert_array_equal((flags[0] >> 1) & 1, a.clipped[:n])
    assert not np.any(flags & 4)
    ref = a.value[:n]
    assert np.mean(value[0] == ref) >= 0.99
    np.testing.assert_allclose(value[0], ref, rtol=1e-14, atol=0)


def test_certified_clearance_matches_reference(attempts, meshes):
    for a in attempts:
        got = O.cert_distance(_geometry(meshes, a), a.point)[0]
        assert got == a.stages[0]["center_clearance_lower"], a.source


def test_distance_is_the_box_union_minimum(golden):
    """Brute force over every triangle's rounded-down box distance (numpy,
    the same roundings) equals the tree's answer; it never exceeds the exact
    distance to the surface."""
    z = np.load(golden / "bunny_1000.npz")
    v, f = z["vertices"], z["faces"]
    m = O.Mesh(v, f)
    g = O.Geometry(O.MESH, mesh=m)
    rng = np.random.default_rng(5)
    pts = v.min(0) + rng.random((64, 3)) * (v.max(0) - v.min(0))
    got = O.cert_distance(g, pts)
    tri = v[f]
    lo, hi = tri.min(axis=1), tri.max(axis=1)
    down = lambda x: np.nextafter(x, -np.inf)
    pos_down = lambda x: np.maximum(0.0, down(x))
    for i, p in enumerate(pts):
        s = np.zeros(len(f))
        for a in range(3):
            gap = np.maximum(0.0, np.maximum(down(lo[:, a] - p[a]), down(p[a] - hi[:, a])))
            s = pos_down(s + pos_down(gap * gap))
        assert got[i] == np.min(pos_down(np.sqrt(s))), i
    exact = np.array([m.closest(p)[0] for p in pts])
    assert np.all(got <= exact)


@pytest.mark.parametrize("index", [0, 7])
def test_philox_streams_estimate_the_same_benefit(attempts, meshes, index):
    """The Philox form (the GPU's production streams) differs from the
    reference's only in its random numbers: on a retained finalist (the
    bunny with the affine source; the housing with its chips and cavities)
    the mean certified value, the mean step count and the capped / clipped
    rates agree with the reference-pinned mt19937_64 form within their
    standard errors, and no containment check fails.  Each sample is on its
    own stream: a counter that ignored the sample index would give one value
    and SE 0."""
    a = attempts[index]
    n = 4096
    geo = _geometry(meshes, a)
    kw = dict(seed=a.seed, walks=n, max_steps=a.max_steps, eps=a.epsilon, radius_cap=a.radius_cap,
              value_cap=a.value_cap, source=O.CERT_AFFINE if a.kind == "bunny" else O.CERT_HOUSING,
              slope=a.slope, threads=O.host_threads())
    mv, ms, mf = O.certify(geo, a.point, **kw)
    pv, ps, pf = O.certify(geo, a.point, rng=O.RNG_PHILOX, **kw)
    assert not np.any(pf & 4) and not np.any(mf & 4)
    for x, y in ((mv[0], pv[0]), (ms[0].astype(float), ps[0].astype(float)),
                 ((mf[0] & 1).astype(float), (pf[0] & 1).astype(float)),
                 (((mf[0] >> 1) & 1).astype(float), ((pf[0] >> 1) & 1).astype(float))):
        se = np.hypot(np.std(x), np.std(y)) / np.sqrt(n)
        assert abs(np.mean(x) - np.mean(y)) <= 4 * se + 1e-15, (np.mean(x), np.mean(y), se)
    assert 0.7 < np.std(pv[0]) / np.std(mv[0]) < 1.4
    assert len(np.unique(pv[0])) > n // 2


def test_philox_streams_are_keyed_by_index(attempts, meshes):
    """A Philox sample depends on (seed, site, sample) only: a prefix of the
    samples, any thread count and a site offset give the same records."""
    a = attempts[0]
    geo = _geometry(meshes, a)
    kw = dict(seed=a.seed, max_steps=a.max_steps, eps=a.epsilon, radius_cap=a.radius_cap, value_cap=a.value_cap,
              source=O.CERT_AFFINE, slope=a.slope, rng=O.RNG_PHILOX)
    pts = np.vstack([a.point, a.point + [0.01, 0.0, 0.0]])
    full = O.certify(geo, pts, walks=256, threads=8, **kw)
    head = O.certify(geo, pts, walks=64, threads=1, **kw)
    for x, y in zip(full, head):
        np.testing.assert_array_equal(x[:, :64], y)
    second = O.certify(geo, pts[1:], walks=64, site_offset=1, **kw)
    for x, y in zip(head, second):
        np.testing.assert_array_equal(x[1:], y)
