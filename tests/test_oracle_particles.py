"""Pins the oracle's stencils, deposits, pushes, reductions, diagnostics and
quiet start (A2-A4, A6-A13)."""
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.special


def test_ngp_cell_hand_cases(orc):
    N = 128
    assert orc.ngp_cell(0.0, N) == N - 1          # round(0) = 0 -> mod1 -> N
    assert orc.ngp_cell(1.0, N) == N - 1
    assert orc.ngp_cell(1 / N, N) == 0
    assert orc.ngp_cell(2.5 / N, N) == 1          # round(2.5) = 2 (ties to even)
    assert orc.ngp_cell(3.5 / N, N) == 3          # round(3.5) = 4
    assert orc.ngp_cell(0.999, N) == N - 1


@pytest.mark.parametrize("R", [6, 7])
def test_gauss_stencil_against_scipy_and_unity(orc, R):
    N = 64
    rng = np.random.default_rng(R)
    for c in list(rng.uniform(-0.05, 1.05, 200)) + [0.0, 0.5 / N, 1.0]:
        cells, wt = orc.gauss_stencil(N, R, c)
        i0 = int(np.rint(c * N))
        i = np.arange(-R, R + 1) + i0
        ref = 0.5 * (scipy.special.erf(((i + 0.5) / N - c) * N) - scipy.special.erf(((i - 0.5) / N - c) * N))
        assert np.max(np.abs(wt - ref)) < 1e-15
        assert np.array_equal(cells, np.mod(i - 1, N))
        assert abs(wt.sum() - 1.0) < 4e-15          # erfc(7.5) ~ 3e-26
        assert np.all(wt >= 0)


def test_gauss_stencil_symmetries(orc):
    N = 64
    for i0 in [0, 3, 40]:
        _, w = orc.gauss_stencil(N, 7, i0 / N)    # centred on a cell: mirror symmetric
        assert np.array_equal(w, w[::-1])
    rng = np.random.default_rng(11)
    for d in rng.uniform(-0.49, 0.49, 50):
        # translating by a whole number of cells keeps the weights when c*N is exact
        c = (10 + d) / N
        c2 = (23 + d) / N
        _, w1 = orc.gauss_stencil(N, 7, c)
        _, w2 = orc.gauss_stencil(N, 7, c2)
        assert np.max(np.abs(w1 - w2)) < 1e-15
        # mirror: delta -> -delta reverses the weights
        _, w3 = orc.gauss_stencil(N, 7, (10 - d) / N)
        assert np.max(np.abs(w1 - w3[::-1])) < 1e-15


def test_gauss_stencil_sixteen_erf_form_is_bit_exact(orc):
    # adjacent cells share an edge with bit-identical erf argument, so the 16-erf
    # evaluation equals the scripts' 30-erf evaluation bitwise
    N = 4096
    rng = np.random.default_rng(5)
    for c in rng.uniform(-0.01, 1.01, 300):
        _, w = orc.gauss_stencil(N, 7, c)
        i0 = int(np.rint(c * N))
        edges = [orc.erf((((i0 + j) - 0.5) / N - c) * N) / 2.0 for j in range(-7, 9)]
        w16 = np.array([edges[k + 1] - edges[k] for k in range(15)])
        assert np.array_equal(w, w16)


def test_fixed_bits_bound_all_configs(orc):
    for P, w in [(8192, 3.125), (2048, 32 * np.pi ** 2 / 3 / 2048 * 64), (2 ** 24, 400 / 4096), (16, 400.0)]:
        q = orc.fxp_frac_bits(P, w)
        assert P * w * 2.0 ** q < 2.0 ** 126


def test_ngp_deposit_is_count_times_w(orc):
    N, P = 128, 8192
    rng = np.random.default_rng(0)
    x = rng.random(P)
    w = 200 / P * N
    rho = orc.deposit(N, 0, x, w=w)
    counts = np.bincount([orc.ngp_cell(v, N) for v in x], minlength=N)
    assert np.array_equal(rho, counts * w)        # 3.125 is exact: sums are exact
    assert np.array_equal(rho, orc.deposit(N, 0, x, w=w, mode=1))


@pytest.mark.parametrize("R", [6, 7])
def test_gauss_deposit_exact_sum_rounded_once(orc, R):
    N, P = 16, 40
    rng = np.random.default_rng(R + 100)
    x = rng.uniform(-0.02, 1.02, P)
    X = rng.random(P)
    w = 0.37
    rho = orc.deposit(N, R, x, X, w=w)
    acc = [Fraction(0)] * N
    for j in range(P):
        cells, wt = orc.gauss_stencil(N, R, (x[j] + X[j]) / 2.0)
        for cl, wv in zip(cells, wt):
            acc[cl] += Fraction(float(wv * w))      # product rounded as in k[2]*w
    assert np.array_equal(rho, np.array([float(a) for a in acc]))
    seq = orc.deposit(N, R, x, X, w=w, mode=1)
    # the scripts' sequential order differs by at most (n-1) eps sum|a| per cell
    assert np.max(np.abs(seq - rho) / (P * (2 * R + 1) * 2.3e-16 * w)) <= 1.0
    assert abs(rho.sum() - P * w) < 1e-12 * P * w


def test_midpoint_push_bruteforce_and_free_streaming(orc):
    N, R, P = 32, 7, 12
    rng = np.random.default_rng(9)
    X = rng.random(P)
    V = rng.standard_normal(P)
    dt = 1 / (6 * N)
    v = V + 0.01 * rng.standard_normal(P)
    x = orc.midpoint_position(X, V, v, dt)
    assert np.max(np.abs(x - (X + (v + V) / 2 * dt))) < 1e-16
    E = rng.standard_normal(N)
    got = orc.push_midpoint(N, R, X, V, x, E, dt)
    for j in range(P):
        c = (x[j] + X[j]) / 2
        i0 = int(np.rint(c * N))
        i = np.arange(-R, R + 1) + i0
        wt = 0.5 * (scipy.special.erf((i + 0.5 - c * N)) - scipy.special.erf((i - 0.5 - c * N)))
        ref = V[j] + np.sum(E[np.mod(i - 1, N)] * wt) * dt
        assert abs(got[j] - ref) < 1e-14
    # E = 0: velocities stay, positions stream freely
    assert np.array_equal(orc.push_midpoint(N, R, X, V, x, np.zeros(N), dt), V)


def test_leapfrog_pieces(orc):
    N, P = 128, 64
    rng = np.random.default_rng(4)
    x = rng.random(P)
    v = np.where(np.arange(P) < P // 2, -1.0, 1.0)
    dt = 1 / (4 * N)
    x2 = orc.drift_half(x, v, dt)
    assert np.all((x2 >= 0) & (x2 <= 1))
    assert np.max(np.abs(np.mod(x + v * dt / 2 - x2 + 0.5, 1) - 0.5)) < 1e-16
    # E = 0: kick leaves v unchanged (both shapes)
    assert np.array_equal(orc.kick(N, 0, v, x2, np.zeros(N), dt), v)
    assert np.array_equal(orc.kick(N, 6, v, x2, np.zeros(N), dt), v)
    # uniform E: NGP kick adds E*dt exactly; Gaussian adds E*dt*sum(wt)
    E = np.full(N, 0.5)
    assert np.array_equal(orc.kick(N, 0, v, x2, E, dt), v + 0.5 * dt)
    assert np.max(np.abs(orc.kick(N, 6, v, x2, E, dt) - (v + 0.5 * dt))) < 1e-17
    w = orc.wrap(np.array([-1e-20, 1.0, 1.5, -0.25, 0.3]))
    assert list(w) == [1.0, 0.0, 0.5, 0.75, 0.3]


def test_tree_sum_and_isapprox(orc):
    rng = np.random.default_rng(8)
    for n in [1, 7, 256, 4096, 4097, 100000]:
        x = rng.standard_normal(n)
        s = orc.tree_sum(x)
        exact = float(sum(Fraction(v) for v in x)) if n <= 5000 else math.fsum(x)
        assert abs(s - exact) <= 1e-15 * np.sum(np.abs(x)) * np.log2(n + 1)
    x = np.array([3.0, 4.0])
    assert orc.norm2(x) == 5.0
    assert orc.norm2(np.array([1e-200, 1e-200])) == pytest.approx(math.sqrt(2) * 1e-200, rel=1e-15)
    E = rng.standard_normal(64)
    assert not orc.isapprox(np.full(64, np.nan), E, 1e-8)   # F .*= NaN: first sweep always runs
    assert orc.isapprox(E, E, 0.0)
    F = E * (1 + 1e-9)
    assert orc.isapprox(F, E, 1e-8) and not orc.isapprox(F, E, 1e-10)
    assert not orc.isapprox(np.zeros(64), E, 1e-8)


def test_diagnostics_closed_forms(orc):
    N, P, W = 64, 2048, 32 * np.pi ** 2 / 3
    v = np.where(np.arange(P) < P // 2, -1.0, 1.0)
    D = orc.diagnostics(np.zeros(N), v, W)
    assert D[0] == 0.0 and D[1] == 1.0 and D[2] == 1.0 and D[3] == 0.0
    E = np.full(N, 0.5)
    D = orc.diagnostics(E, v * 2, W)
    assert D[0] == pytest.approx(0.25 / W, rel=1e-15)
    assert D[1] == pytest.approx(4.0, rel=1e-15)


def test_quiet_start_bit_pattern(orc):
    P = 2048
    x, v = orc.quiet_start(P)
    for i in range(P):
        r = int(format(i, "064b")[::-1], 2)          # bitreverse of the 64-bit integer
        ref = Fraction((r - 2 ** 64 if r >= 2 ** 63 else r) + 2 ** 63, 2 ** 64)
        assert Fraction(x[i]) == ref
    assert len(set(x)) == P
    assert np.all(v[: P // 2] == -1) and np.all(v[P // 2:] == 1)
    # each beam is a perfect lattice of spacing 2/P, offset from the other by 1/P
    a, b = np.sort(x[: P // 2]), np.sort(x[P // 2:])
    assert np.all(np.diff(a) == 2 / P) and np.all(np.diff(b) == 2 / P)
    assert abs(b[0] - a[0]) == 1 / P
