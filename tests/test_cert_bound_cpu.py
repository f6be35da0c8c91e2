"""The production certified sampler's distance upper bound (twos_cert.hip,
dist_upper_fast): U = sqrt(dx^2 + dy^2 + dz^2) (1 + 2^-49) + 2^-500 with
round-to-nearest binary64 operations in that order.  It must be at least
|a - b| for every pair of binary64 points, also when the device's square root
is one ulp low.  Checked here in exact rational arithmetic (U^2 >= the exact
squared distance of the two binary64 points) on random, nearly equal, huge,
tiny and subnormal differences.  A bound without the scale factor fails."""

from fractions import Fraction

import numpy as np


def _upper(a, b, sqrt_ulps_low=0, scale=1.0 + 2.0**-49, add=2.0**-500):
    dx, dy, dz = a[0] - b[0], a[1] - b[1], a[2] - b[2]
    s = dx * dx + dy * dy + dz * dz
    r = np.sqrt(np.float64(s))
    for _ in range(sqrt_ulps_low):
        r = np.nextafter(r, 0.0)
    return float(r * np.float64(scale) + np.float64(add))


def _exact_sq(a, b):
    return sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(a, b))


def _pairs():
    rng = np.random.default_rng(17)
    out = []
    for _ in range(400):
        a = rng.uniform(-2, 2, 3)
        out.append((a, a + rng.normal(size=3) * 10.0 ** rng.uniform(-17, 0)))
    for _ in range(100):
        a = rng.uniform(-1, 1, 3)
        out.append((a, np.nextafter(a, a + 1.0)))  # one ulp apart per axis
    for scale in (1e-160, 1e-200, 1e-300, 1e150, 1e300):
        a = rng.uniform(-1, 1, 3) * scale
        out.append((a, a * (1 + 1e-15)))
        out.append((a, -a))
    out.append((np.array([5e-324, 0.0, 0.0]), np.zeros(3)))  # the smallest subnormal
    return out


def test_upper_bound_is_rigorous():
    for a, b in _pairs():
        exact = _exact_sq(a, b)
        for low in (0, 1):
            u = _upper(a, b, sqrt_ulps_low=low)
            # An overflowing square gives U = inf, a (useless but valid) bound.
            assert u >= 0.0 and (np.isinf(u) or Fraction(u) ** 2 >= exact), (a, b, low)


def test_the_scale_factor_is_needed():
    """Without the 1 + 2^-49 factor, some pair's computed distance falls below
    the exact one (the rounding the factor covers does occur)."""
    short = 0
    for a, b in _pairs():
        u = _upper(a, b, sqrt_ulps_low=1, scale=1.0, add=0.0)
        if np.isfinite(u) and Fraction(u) ** 2 < _exact_sq(a, b):
            short += 1
    assert short > 0
