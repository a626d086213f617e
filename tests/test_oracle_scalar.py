"""Pins the oracle's scalar building blocks: Julia semantics (A1), the
OpenLibm-restated exp/erf (N2, A7) and the fixed-point accumulator."""
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.special

from conftest import ulp_diff


def test_julia_mod_float(orc):
    # Julia mod(x, 1): result in [0, 1], mod(-tiny, 1) rounds to 1.0
    assert orc.jl_mod1f(-1e-20) == 1.0
    assert orc.jl_mod1f(0.25) == 0.25
    assert orc.jl_mod1f(1.0) == 0.0 and math.copysign(1, orc.jl_mod1f(-1.0)) == 1.0
    assert orc.jl_mod1f(-0.25) == 0.75
    assert orc.jl_mod1f(1.75) == 0.75
    assert orc.jl_mod1f(-1.5) == 0.5
    assert orc.jl_mod1f(3.0 + 2 ** -50) == 2 ** -50 + 0.0


def test_julia_mod1_int(orc):
    assert orc.jl_mod1i(0, 8) == 8
    assert orc.jl_mod1i(8, 8) == 8
    assert orc.jl_mod1i(1, 8) == 1
    assert orc.jl_mod1i(-1, 8) == 7
    assert orc.jl_mod1i(-8, 8) == 8
    assert orc.jl_mod1i(17, 8) == 1


def test_julia_round_ties_to_even(orc):
    assert orc.jl_round(2.5) == 2.0
    assert orc.jl_round(3.5) == 4.0
    assert orc.jl_round(-2.5) == -2.0
    assert orc.jl_round(0.49999999999999994) == 0.0


def _exp_points():
    rng = np.random.default_rng(1)
    pts = [0.0, -0.0, 1e-300, -1e-300, 2 ** -29, -(2 ** -29), 0.3, -0.3, 0.5, -0.5,
           0.6931471805599453, 1.0, -1.0, 1.0397, -1.0397, 2.0, -2.0, 10.0, -10.0,
           -0.5625, -36.5625, 700.0, -700.0, 709.7, -745.0]
    pts += list(rng.uniform(-0.35, 0.35, 2000))       # k == 0 branch
    pts += list(rng.uniform(0.35, 1.04, 2000))        # |x| < 1.5 ln2 branch
    pts += list(-rng.uniform(0.35, 1.04, 2000))
    pts += list(rng.uniform(-40.0, 40.0, 4000))       # general reduction
    pts += list(rng.uniform(-708.0, 709.0, 2000))
    return np.array(pts)


def test_exp_matches_glibc_within_one_ulp(orc):
    xs = _exp_points()
    mine = orc.exp_array(xs)
    ref = np.array([math.exp(v) for v in xs])
    assert np.max(ulp_diff(mine, ref)) <= 1
    assert orc.exp(float("inf")) == float("inf") and orc.exp(float("-inf")) == 0.0
    assert math.isnan(orc.exp(float("nan")))
    assert orc.exp(800.0) == float("inf") and orc.exp(-800.0) == 0.0


def _erf_points():
    rng = np.random.default_rng(2)
    pts = [0.0, 1e-310, 2 ** -29, 0.84375, 0.8437499999999999, 1.25, 1.2499999999999998,
           1 / 0.35, 2.857142857142857, 5.999999999999999, 6.0, 6.5, 7.5, 28.0]
    for lo, hi in [(0, 2 ** -28), (2 ** -28, 0.84375), (0.84375, 1.25), (1.25, 1 / 0.35),
                   (1 / 0.35, 6.0), (6.0, 30.0)]:
        pts += list(rng.uniform(lo, hi, 3000))
    # the arguments the +-7 stencil produces: j + 0.5 - delta, delta in [-0.5, 0.5]
    pts += list((np.arange(-7, 9)[:, None] - 0.5 - rng.uniform(-0.5, 0.5, (16, 300))).ravel())
    pts = np.array(pts)
    return np.concatenate([pts, -pts])


def test_erf_matches_glibc_within_one_ulp(orc):
    xs = _erf_points()
    mine = orc.erf_array(xs)
    ref = np.array([math.erf(v) for v in xs])
    assert np.max(ulp_diff(mine, ref)) <= 1


def test_erf_matches_scipy(orc):
    xs = _erf_points()
    mine = orc.erf_array(xs)
    ref = scipy.special.erf(xs)
    # scipy's Cephes erf is itself a few ulp from the true value
    assert np.max(ulp_diff(mine, ref)) <= 3


def test_erf_closed_forms(orc):
    assert orc.erf(0.0) == 0.0
    assert math.copysign(1.0, orc.erf(-0.0)) == -1.0
    for v in [6.0, 6.25, 7.5, 100.0, float("inf")]:
        assert orc.erf(v) == 1.0 and orc.erf(-v) == -1.0
    assert math.isnan(orc.erf(float("nan")))
    xs = _erf_points()
    pos = orc.erf_array(np.abs(xs))
    neg = orc.erf_array(-np.abs(xs))
    assert np.array_equal(pos, -neg)          # odd symmetry, bitwise
    s = np.sort(np.abs(xs))
    assert np.all(np.diff(orc.erf_array(s)) >= 0)  # monotone
    # erf(x) = 1 exactly once erfc(x) < 2^-54 (x > 5.921...), as the +-7
    # stencil's outer cells rely on
    assert orc.erf(5.93) == 1.0 and orc.erf(5.9) < 1.0


def test_fixed_point_conversion_exact(orc):
    rng = np.random.default_rng(3)
    q = orc.fxp_frac_bits(2 ** 24, 400.0 / 4096)
    assert 2 ** 24 * (400.0 / 4096) * 2.0 ** q < 2.0 ** 126
    for a in list(rng.uniform(0, 1, 200)) + [0.0, 5e-324, 1e-300, 2.0 ** -q, 1.5 * 2.0 ** -q,
                                            2.5 * 2.0 ** -q, 0.5 * 2.0 ** -q, 3.0]:
        Iv = orc.fxp_from_double(a, q)
        exact = Fraction(a) * 2 ** q
        # round half to even
        assert Iv == round(exact)
    # sums of many contributions convert with correct rounding
    vals = rng.uniform(0, 0.1, 5000) * rng.choice([1, 1e-9, 1e-17], 5000)
    S = sum(orc.fxp_from_double(v, q) for v in vals)
    got = orc.fxp_to_double(S, q)
    assert got == float(Fraction(S, 2 ** q))
    assert got == float(sum(Fraction(v) for v in vals))  # exact sum, rounded once


def test_fixed_point_edges_are_defined(orc):
    """Rounding onto the 2^-q grid at the edges of its domain: Inf and NaN
    add nothing, a value at or beyond 2^128 grid units saturates, the
    largest representable shift is exact, and ties round to even."""
    q = 64
    assert orc.fxp_from_double(float("nan"), q) == 0
    assert orc.fxp_from_double(float("inf"), q) == 0
    assert orc.fxp_from_double(2.0 ** 70, q) == (1 << 128) - 1          # 2^134 units
    big = (2 ** 53 - 1) * 2.0 ** 10                                       # mantissa shifted by 74
    assert orc.fxp_from_double(big, q) == (2 ** 53 - 1) << 74
    assert orc.fxp_from_double(2.5 * 2.0 ** -q, q) == 2                   # half to even
    assert orc.fxp_from_double(3.5 * 2.0 ** -q, q) == 4
    assert orc.fxp_from_double(2.0 ** -(q + 2), q) == 0                   # below half a unit
