"""The mixed push's +-5 cell window against the configured +-7 Gaussian
stencil (src/GaussianFixedPoint.jl:4-5, src/GaussianFixedPointQuiet.jl:5-6).

1. The FP64 weights of cells |j| = 6, 7 are below half a deposit quantum
   (2^-31 of a particle's charge) for every sub-cell offset, so the +-7
   deposit rounded to the mixed path's 2^-30 grid is the +-5 deposit.
2. The FP32 gather over +-5 -- the quanta polynomials of
   csrc/weights32.h evaluated with FP32 fused multiply-adds, the field
   scaled and rounded to FP32 once per cell as the solve does -- agrees with
   the oracle's FP64 +-7 gather to within the FP32 rounding bound of the
   sum plus the stated fit and evaluation errors of the weights.
"""
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "..", "particleincellcodegolf.jl_amd", "csrc", "weights32.h")
U = 2.0 ** -24   # FP32 unit roundoff


def quanta_polys():
    """{name: coefficients from the highest degree down} of the q* functions."""
    src = open(WEIGHTS).read()
    out = {}
    for m in re.finditer(r"float (q[AB]\d)\(float u\) \{(.*?)return a;", src, re.S):
        body = m.group(2)
        c0 = re.search(r"float a = ([-0-9.e+]+)f;", body).group(1)
        rest = re.findall(r"a = fmaf\(a, u, ([-0-9.e+]+)f\);", body)
        out[m.group(1)] = [np.float32(float(c0))] + [np.float32(float(c)) for c in rest]
    return out


def fmaf(a, b, c):
    # a*b is exact in float64 for float32 operands; one rounding to float32
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def horner(coefs, u):
    a = coefs[0]
    for c in coefs[1:]:
        a = fmaf(a, u, c)
    return a


def weights_q(P, d):
    """Window slots 0..10 (cells i0-5 .. i0+5) in 2^-30 units, as stencil32."""
    d = np.float32(d)
    u = np.float32(d * d)
    w = [np.float32(0)] * 11
    w[5] = horner(P["qA0"], u)
    for j in range(1, 6):
        a = horner(P[f"qA{j}"], u)
        b = np.float32(d * horner(P[f"qB{j}"], u))
        w[5 + j] = np.float32(a + b)
        w[5 - j] = np.float32(a - b)
    return w


def test_weights_beyond_five_cells_quantise_to_zero(orc):
    N = 64
    for d in np.linspace(-0.5, 0.5, 2001):
        c = (10 + d) / N          # centre d cells from the centre of cell 10
        cells, wt = orc.gauss_stencil(N, 7, c)
        far = np.abs(np.arange(-7, 8)) >= 6
        assert np.max(wt[far]) * 2.0 ** 30 < 0.5
        # the +-5 weights carry the particle: 1 - sum below 2e-12
        assert abs(1.0 - np.sum(wt[~far])) < 2e-12


def test_fp32_window_gather_within_rounding_of_fp64_stencil(orc):
    P = quanta_polys()
    assert len(P) == 11
    N = 64
    dt = 1.0 / (6 * N)
    s = dt * dt * 4.0      # the solve's scale of the mixed gather's field
    rng = np.random.default_rng(5)
    worst = 0.0
    for trial in range(400):
        E = rng.standard_normal(N) * 10.0 ** rng.uniform(-3, 3)
        d = rng.uniform(-0.5, 0.5)
        i0 = int(rng.integers(0, N))
        c = (i0 + d) / N
        ref = orc.gather_at(N, 7, E, c)
        w = weights_q(P, d)
        E32 = [np.float32(E[(i0 - 1 + j - 5) % N] * s) for j in range(11)]  # Julia cell i0+j-5 at array i0+j-6
        k = np.float32(0)
        for j in range(11):
            k = fmaf(E32[j], w[j], k)
        got = float(k) / (s * 2.0 ** 30)
        cells, wt = orc.gauss_stencil(N, 7, c)
        absE = np.abs(E[cells])
        # field rounding, the 11-term FMA chain, the weights' FP32 evaluation
        # (<= 3.74e-8) and fit (<= 1.64e-9) errors, the dropped |j| >= 6 tails
        bound = np.sum(absE * np.abs(wt)) * 14 * U + np.sum(absE[1:-1]) * (3.74e-8 + 1.64e-9 + 2 * U) \
            + np.sum(absE[[0, 1, -2, -1]] * wt[[0, 1, -2, -1]])
        assert abs(got - ref) <= bound, (trial, got, ref, bound)
        worst = max(worst, abs(got - ref) / bound)
    print(f"worst error / bound {worst:.3f}")
