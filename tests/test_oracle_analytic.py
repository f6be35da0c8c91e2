"""The analytic source (source_kind 3; the reference's exact_bump_green,
study.cpp:56-76) in the oracle.

No retained output of the analytic-source executable exists, so it is pinned
by what the mathematics fixes:

- the closed form equals a direct quadrature of the integral of the ball's
  Green function G(x, y) = 1/(4 pi |x-y|) - 1/(4 pi r) against the bump, in
  every regime (ball inside the bump, straddling its edge, enclosing it,
  concentric, tangent, disjoint);
- a path from the centre of the unit ball takes one step, so its value is
  that integral exactly and every walk agrees;
- the analytic and the sampled (Green-point) sources estimate the same mean,
  and the analytic one with less variance.
"""

import math

import numpy as np
import pytest
from scipy import integrate

import oracle as O


def _quadrature(d, r, R, A):
    """Integral over the ball |y - x| < r of G(x, y) A (1 - s^2/R^2)_+^2 with
    s = |y - c|, |x - c| = d, in (rho, mu) coordinates about x."""
    bump = lambda s2: A * max(0.0, 1.0 - s2 / (R * R)) ** 2

    def shell(rho):
        if rho == 0.0:
            return 0.0
        g = 1 / (4 * math.pi * rho) - 1 / (4 * math.pi * r)
        s2 = lambda mu: d * d + rho * rho + 2 * d * rho * mu
        pts = []
        if d > 0:
            mu_edge = (R * R - d * d - rho * rho) / (2 * d * rho)
            if -1 < mu_edge < 1:
                pts = [mu_edge]
        inner = integrate.quad(lambda mu: bump(s2(mu)), -1, 1, points=pts or None, epsabs=1e-15, epsrel=1e-13,
                               limit=200)[0]
        return 2 * math.pi * rho * rho * g * inner

    brk = [b for b in (abs(R - d), R + d) if 0 < b < r]
    return integrate.quad(shell, 0, r, points=brk or None, epsabs=1e-15, epsrel=1e-12, limit=400)[0]


@pytest.mark.parametrize("d,r,R,A", [
    (0.02, 0.05, 0.3, 2.0),     # the ball well inside the bump
    (0.25, 0.1, 0.2, 35.0),     # straddling the bump's edge
    (0.05, 0.5, 0.1, 4.0),      # the ball encloses the bump (d + R <= r)
    (0.0, 0.1, 0.2, 1.0),       # concentric, ball inside
    (0.0, 0.3, 0.2, 1.0),       # concentric, ball encloses
    (0.3, 0.2, 0.1 + 1e-9, 3.0),  # nearly tangent from outside
    (0.12, 0.3, 0.2, 1.5),      # the sphere cuts the bump
])
def test_closed_form_equals_quadrature(d, r, R, A):
    exact = _quadrature(d, r, R, A)
    got = O.bump_green_ball([d, 0.0, 0.0], r, [0.0, 0.0, 0.0], R, A)
    assert abs(got - exact) <= 1e-9 * max(abs(exact), 1e-12), (got, exact)
    # Direction-independent: the same d along another axis.
    got2 = O.bump_green_ball([0.3, 0.1, -0.2], r, [0.3, 0.1, -0.2 + d], R, A)
    assert abs(got2 - got) <= 1e-12 * max(abs(got), 1e-12)


def test_disjoint_ball_gives_zero():
    assert O.bump_green_ball([1.0, 0.0, 0.0], 0.2, [0.0, 0.0, 0.0], 0.3, 2.0) == 0.0


def _unit_ball():
    return O.Geometry(O.BALL, ball=(0.0, 0.0, 0.0, 1.0))


ROWS = [(O.TERM_BUMP, 0.2, 0.0, 0.1, 0.3, 0.0, 0.0, 2.0), (O.TERM_BUMP, -0.3, 0.2, 0.0, 0.25, 0.0, 0.0, 5.0)]


def test_one_step_from_the_centre_is_exact():
    side = O.side(terms=ROWS, analytic=True)
    r = O.estimate(_unit_ball(), [[0.0, 0.0, 0.0]], [side], combine=O.SQUARE, rng=O.RNG_PHILOX, seed=1, walks=64,
                   eps=1e-6, max_steps=16, dump=True)
    expect = sum(_quadrature(math.hypot(*row[1:4]), 1.0, row[4], row[7]) for row in ROWS)
    assert np.all(r["walk_steps"][0] == 1)
    assert np.all(r["a"][0] == r["a"][0][0])
    assert abs(r["u"][0] - expect) <= 1e-9 * expect


@pytest.mark.parametrize("rng", [O.RNG_PHILOX, O.RNG_MT_PAIRED])
def test_analytic_and_sampled_sources_agree(rng):
    sites = np.array([[0.1, -0.2, 0.3], [-0.4, 0.1, 0.0], [0.25, 0.25, -0.25]])
    kw = dict(combine=O.PAIRED, rng=rng, seed=7, walks=20000, eps=1e-4, max_steps=4096, threads=8)
    zero = O.side()
    sampled = O.estimate(_unit_ball(), sites, [O.side(terms=ROWS), zero], **kw)
    exact = O.estimate(_unit_ball(), sites, [O.side(terms=ROWS, analytic=True), zero], **kw)
    n = 20000
    # The per-walk spread of each source, from SQUARE runs (same = mean a^2).
    sq = O.estimate(_unit_ball(), sites, [O.side(terms=ROWS)], combine=O.SQUARE, rng=O.RNG_PHILOX, seed=8,
                    walks=n, eps=1e-4, max_steps=4096, threads=8)
    se_sampled = np.sqrt(np.maximum(sq["same"] - sq["u"] ** 2, 0) / n)
    sq_a = O.estimate(_unit_ball(), sites, [O.side(terms=ROWS, analytic=True)], combine=O.SQUARE,
                      rng=O.RNG_PHILOX, seed=8, walks=n, eps=1e-4, max_steps=4096, threads=8)
    se_exact = np.sqrt(np.maximum(sq_a["same"] - sq_a["u"] ** 2, 0) / n)
    assert np.all(np.abs(sampled["u"] - exact["u"]) < 4 * np.hypot(se_sampled, se_exact))
    # Rao-Blackwellisation: the analytic source removes the Green point's noise.
    assert np.all(se_exact < 0.8 * se_sampled)


def test_analytic_rejects_gaussian_rows():
    side = O.side(terms=[(O.TERM_GAUSS, 0, 0, 0, 0.2, 0.2, 0.2, 1.0)], analytic=True)
    with pytest.raises(ValueError):
        O.estimate(_unit_ball(), [[0.0, 0.0, 0.0]], [side], combine=O.SQUARE, walks=8)
