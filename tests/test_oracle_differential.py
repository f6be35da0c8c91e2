"""Differential WoS cavity derivatives (evaluate_differential_site,
benchmarks/compiled_bunny/src/wih_twos.cpp:725-928, as wih_study --diff 1
runs it, benchmarks/submission_study/study.cpp:303-316) in the oracle,
against the reference's retained requests on the branch mesh.

Pins: the request totals of kernel.json (steps = outer + nested steps,
distance queries, parity rays, caps) exactly, which fixes the outer and
nested streams, the offset-safety classification (its parity rays and
queries are counted) and every termination; the derivative columns to
2e-4 of their scale, and most sites to the last bit (near-tie closest
triangles move a few paths by ulps, as in the paired study requests).
"""

import numpy as np
import pytest

import oracle as O
import study_fixtures as SF


@pytest.fixture(scope="module")
def diff(golden):
    v, f, requests = SF.load_diff(golden)
    return O.Mesh(v, f), requests


def _run(mesh, req, threads=8, points=None):
    heat, _ = O.study_terms(req.config)
    g = O.Geometry(O.MESH, mesh=mesh, cavities=req.holes, cavity_radius=req.cavity_radius)
    return O.differential(g, req.points if points is None else points, O.side(terms=heat), seed=req.seed,
                          walks=req.walks, eps=req.epsilon, max_steps=req.max_steps, offset=req.offset,
                          threads=threads)


def _totals(counts):
    c = dict(zip(O.DIFF_COUNTS, counts.sum(axis=0)))
    return {"steps": int(c["outer_steps"] + c["nested_steps"]), "distance_queries": int(c["distance_queries"]),
            "parity_rays": int(c["parity_rays"]), "caps": int(c["limits"])}


@pytest.mark.parametrize("source", ["validation-local-cavity/raw/gradient-0.004-0",
                                    "validation-local-cavity/raw/gradient-0.008-2",
                                    "validation/raw/gradient-0.002-0"])
def test_differential_requests(diff, source):
    mesh, requests = diff
    req = next(r for r in requests if r.source == source)
    values, counts = _run(mesh, req)
    assert _totals(counts) == req.totals
    exact = SF.compare_diff(req, values, rtol=2e-4)
    assert exact.mean() > 0.9, exact.mean()
    assert np.count_nonzero(req.values) > 0


def test_site_streams_are_keyed_by_index(diff):
    """A site's outer and nested streams depend on (seed, repeat, site, walk)
    only: a prefix of the sites and any thread count give the same values."""
    mesh, requests = diff
    req = requests[0]
    full, _ = _run(mesh, SF.DiffRequest(**{**req.__dict__, "walks": 64}), threads=8)
    head, _ = _run(mesh, SF.DiffRequest(**{**req.__dict__, "walks": 64}), threads=1, points=req.points[:20])
    np.testing.assert_array_equal(head, full[:20])


def test_offset_safety_rejects_overlapping_cavities(diff):
    """Two cavities closer than twice the radius: hits near their seam circle
    are classified unsafe and contribute nothing (classify_offset_safety,
    wih_twos.cpp:764-833)."""
    mesh, requests = diff
    req = requests[0]
    c = req.holes[0]
    holes = np.array([c, c + [0.6 * req.cavity_radius, 0.0, 0.0]])
    heat, _ = O.study_terms(req.config)
    g = O.Geometry(O.MESH, mesh=mesh, cavities=holes, cavity_radius=req.cavity_radius)
    pts = req.points[:24]
    _, counts = O.differential(g, pts, O.side(terms=heat), seed=3, walks=128, eps=req.epsilon,
                               max_steps=req.max_steps, offset=req.offset, threads=8)
    c = dict(zip(O.DIFF_COUNTS, counts.sum(axis=0)))
    assert c["cavity_terminations"] > 0 and c["near_seam"] > 0
    assert c["unsafe"] >= c["near_seam"]


def _run_philox(mesh, req, walks, points=None, threads=8, seed=None):
    heat, _ = O.study_terms(req.config)
    g = O.Geometry(O.MESH, mesh=mesh, cavities=req.holes, cavity_radius=req.cavity_radius)
    return O.differential(g, req.points if points is None else points, O.side(terms=heat),
                          seed=req.seed if seed is None else seed, walks=walks, eps=req.epsilon,
                          max_steps=req.max_steps, offset=req.offset, threads=threads, rng=O.RNG_PHILOX)


def test_philox_streams_estimate_the_same_derivatives(diff):
    """The Philox form (the GPU's production streams) differs from the
    reference's only in its random numbers: on a retained request it gives
    the same derivatives as the reference-pinned mt19937_64 form within
    their standard errors, the same step and termination rates, and
    standard errors of the same size over the request (each sample on its
    own stream: a counter that ignored the sample index would repeat one path
    and leave SE 0)."""
    mesh, requests = diff
    req = next(r for r in requests if r.source == "validation-local-cavity/raw/gradient-0.004-0")
    mt, cm = _run(mesh, SF.DiffRequest(**{**req.__dict__, "walks": 512}))
    ph, cp = _run_philox(mesh, req, 512)
    names = O.DIFF_COUNTS
    tm, tp = dict(zip(names, cm.sum(axis=0))), dict(zip(names, cp.sum(axis=0)))
    for k in ("outer_steps", "nested_steps", "distance_queries", "cavity_terminations", "parity_rays"):
        assert abs(tp[k] - tm[k]) <= 0.03 * tm[k], (k, tp[k], tm[k])
    for col, se in ((0, 3), (1, 4), (2, 5), (6, 7)):
        s = np.hypot(mt[..., se], ph[..., se])
        ok = s > 0
        assert ok.sum() >= 100, ok.sum()
        z = (mt[..., col] - ph[..., col])[ok] / s[ok]
        assert np.max(np.abs(z)) < 4.0 and abs(np.mean(z)) < 0.35, (col, np.max(np.abs(z)), np.mean(z))
        # Few samples per site end safely on the cavity, so per-site SEs are
        # noisy: compare their root mean square over the request.
        ratio = np.sqrt(np.sum(ph[..., se] ** 2) / np.sum(mt[..., se] ** 2))
        assert 0.5 < ratio < 2.0, (col, ratio)


def test_philox_streams_are_keyed_by_index(diff):
    """A Philox sample depends on (seed, repeat, site, sample) only: a prefix
    of the sites and any thread count give the same values, and another seed
    other values."""
    mesh, requests = diff
    req = requests[0]
    full, fc = _run_philox(mesh, req, 64, threads=8)
    head, hc = _run_philox(mesh, req, 64, threads=1, points=req.points[:20])
    np.testing.assert_array_equal(head, full[:20])
    np.testing.assert_array_equal(hc, fc[:20])
    other, _ = _run_philox(mesh, req, 64, points=req.points[:20], seed=req.seed + 1)
    assert not np.array_equal(other, head)
