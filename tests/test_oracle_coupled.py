"""The coupled finite-cavity difference (study.cpp --fd, study.cpp:317-332) in
the oracle.

No retained output of an --fd run exists in the reference (every recorded
request passes --fd 0), so the estimator is pinned by what the mathematics
fixes:

- coupling: a cavity no path can reach leaves the second path identical to
  the first, so every sample is exactly zero (independent streams for the two
  paths would not be);
- a closed form: in the unit ball with a constant source, a concentric cavity
  and a concentric probe ball, J with and without the cavity are radial
  integrals of the two exact solutions, so the benefit per capacity
  (J - J_cavity) / (4 pi r) is known; a dropped weight factor, a uniform
  instead of cube-root radius, a wrong capacity or sign fails it;
- the reference's mean and SE formula (study.cpp:329-331) against numpy on
  the per-walk samples.
"""

import numpy as np
import pytest
from scipy import integrate

import oracle as O


def _ball(cavity_radius):
    return O.Geometry(O.BALL, ball=(0.0, 0.0, 0.0, 1.0), cavity_radius=cavity_radius)


def _probe_side(center, radius):
    mass = 32 * np.pi * radius**3 / 105
    return O.side(terms=[(O.TERM_UNIT, *center, radius, mass, 0.0, 0.0)])


CONST = O.side(source=(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0))


@pytest.mark.parametrize("rng", [O.RNG_MT_FD, O.RNG_PHILOX])
def test_cavity_out_of_reach_gives_exact_zero(rng):
    """Every path stays in the unit ball, so a cavity centred 4 away never
    wins the distance minimum: both paths coincide step for step."""
    sites = np.array([[5.0, 0.0, 0.0], [0.0, -4.0, 0.5]])
    r = O.estimate(_ball(0.1), sites, [CONST, _probe_side((0.1, 0.2, 0.0), 0.3)], combine=O.COUPLED, rng=rng,
                   seed=3, walks=64, eps=1e-4, max_steps=4096, dump=True)
    assert np.all(r["a"] == 0.0)
    np.testing.assert_array_equal(r["score"], [0.0, 0.0])
    np.testing.assert_array_equal(r["score_se"], [0.0, 0.0])
    # Two identical paths per walk: an even step count per walk.
    assert np.all(r["walk_steps"] % 2 == 0) and np.all(r["walk_steps"] > 0)


def _concentric_exact(rc, rs):
    """(J0 - J1) / (4 pi rc) for -Lap u = 1 in B(0,1), u = 0 on the boundary,
    the cavity B(0, rc) and q the unit-mass bump on B(0, rs)."""
    u0 = lambda p: (1 - p * p) / 6
    b = -rc * (1 + rc) / 6
    a = 1 / 6 - b
    u1 = lambda p: np.where(p < rc, 0.0, -p * p / 6 + a + b / np.maximum(p, 1e-300))
    q = lambda p: (1 - (p / rs) ** 2) ** 2 * 105 / (32 * np.pi * rs**3)
    diff = integrate.quad(lambda p: 4 * np.pi * p * p * q(p) * (u0(p) - u1(p)), 0, rs, points=[rc], limit=200)[0]
    return diff / (4 * np.pi * rc)


def test_concentric_exact_solution_checks():
    # The cavity solution vanishes on both spheres and solves -Lap u = 1.
    rc = 0.2
    b = -rc * (1 + rc) / 6
    a = 1 / 6 - b
    u1 = lambda p: -p * p / 6 + a + b / p
    assert abs(u1(1.0)) < 1e-15 and abs(u1(rc)) < 1e-15
    p, h = 0.5, 1e-4
    lap = (u1(p + h) - 2 * u1(p) + u1(p - h)) / h**2 + 2 / p * (u1(p + h) - u1(p - h)) / (2 * h)
    assert abs(lap + 1) < 1e-5


@pytest.mark.parametrize("rng", [O.RNG_MT_FD, O.RNG_PHILOX])
def test_concentric_closed_form(rng):
    rc, rs = 0.2, 0.6
    exact = _concentric_exact(rc, rs)
    r = O.estimate(_ball(rc), [[0.0, 0.0, 0.0]], [CONST, _probe_side((0.0, 0.0, 0.0), rs)], combine=O.COUPLED,
                   rng=rng, seed=11, walks=40000, eps=1e-5, max_steps=8192, threads=8)
    score, se = r["score"][0], r["score_se"][0]
    assert se < 0.02 * exact
    assert abs(score - exact) < 4 * se, (score, exact, se)
    # A probe drawn with a uniform radius (rho = R U instead of R U^(1/3))
    # would weight the centre, where the difference is largest, more.
    uniform_radius = integrate.quad(
        lambda p: (1 - (p / rs) ** 2) ** 2 * 35 / 8 / rs * ((1 - p * p) / 6 - (0.0 if p < rc else
                  -p * p / 6 + 1 / 6 + rc * (1 + rc) / 6 - rc * (1 + rc) / 6 / p)), 0, rs, points=[rc])[0]
    assert abs(uniform_radius / (4 * np.pi * rc) - exact) > 10 * se


def test_mean_and_se_follow_the_reference_formula():
    sides = [CONST, _probe_side((0.1, 0.0, 0.0), 0.4)]
    r = O.estimate(_ball(0.15), [[0.2, 0.1, 0.0], [-0.3, 0.0, 0.2]], sides, combine=O.COUPLED, rng=O.RNG_MT_FD,
                   seed=5, walks=257, eps=1e-4, max_steps=4096, dump=True)
    for i in range(2):
        d = r["a"][i]
        assert np.all(d == r["b"][i])
        np.testing.assert_allclose(r["score"][i], d.mean(), rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(r["score_se"][i], d.std(ddof=1) / np.sqrt(len(d)), rtol=1e-9)
        assert r["walks"][i] == 257 and r["steps"][i] == r["walk_steps"][i].sum()


def test_coupled_needs_a_unit_probe_and_a_cavity_radius():
    with pytest.raises(ValueError):
        O.estimate(_ball(0.0), [[0.0, 0.0, 0.0]], [CONST, _probe_side((0, 0, 0), 0.3)], combine=O.COUPLED,
                   rng=O.RNG_MT_FD, walks=8)
    with pytest.raises(ValueError):
        O.estimate(_ball(0.1), [[0.0, 0.0, 0.0]], [CONST, CONST], combine=O.COUPLED, rng=O.RNG_MT_FD, walks=8)
    with pytest.raises(ValueError):  # the fd stream layout belongs to COUPLED
        O.estimate(_ball(0.1), [[0.0, 0.0, 0.0]], [CONST, CONST], combine=O.SQUARE, rng=O.RNG_MT_FD, walks=8)
