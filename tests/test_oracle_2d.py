"""Pins the 2D3V oracle (oracle/pic2d.c) against what the mathematics and the
reference's expressions fix (src/Electrostatic2D3V.jl): hand cases of unimod
and of the area weights g(), exact rational deposits, numpy's FFT and the O(N^4)
DFT, single-mode Poisson closed forms, the textbook vector Boris rotation, and
whole-run plasma physics (plasma and upper-hybrid frequencies, x momentum)."""
from fractions import Fraction

import numpy as np
import pytest


def test_unimod_hand_cases(orc):
    L = orc.lib()
    assert L.orc2d_unimod(0.0, 1.0) == 1.0          # 0 -> n: positions live in (0, 1]
    assert L.orc2d_unimod(1.0, 1.0) == 1.0
    assert L.orc2d_unimod(1.25, 1.0) == 0.25
    assert L.orc2d_unimod(-0.25, 1.0) == 0.75
    assert L.orc2d_unimodi(0, 8) == 8
    assert L.orc2d_unimodi(9, 8) == 1
    assert L.orc2d_unimodi(-1, 8) == 7
    assert L.orc2d_unimodi(5, 8) == 5


def test_area_weights_hand_cases(orc):
    # z = 5/16, NZ = 8: zN = 2.5, i = ceil = 3 (1-based), r = 0.5
    assert orc.area_weights(5 / 16, 8) == ((2, 0.5), (3, 0.5))
    # z = 0.3: i = 3, r = 3 - 2.4: cell i gets 1 - r, cell i+1 gets r (the reference pairing)
    (c0, w0), (c1, w1) = orc.area_weights(0.3, 8)
    assert (c0, c1) == (2, 3)
    assert w1 == 3.0 - 0.3 * 8 and w0 == 1.0 - w1
    # textbook variant pairs i with i-1
    (c0, _), (c1, _) = orc.area_weights(0.3, 8, textbook=1)
    assert (c0, c1) == (2, 1)
    # z = 1 (== 0 of the periodic box): zN = NZ, r = 0, all weight on cell NZ, partner wraps to cell 1
    assert orc.area_weights(1.0, 8) == ((7, 1.0), (0, 0.0))
    assert orc.area_weights(1.0, 8, textbook=1) == ((7, 1.0), (6, 0.0))
    # tiny z: ceil = 1, r ~ 1: the partner (cell 2) takes almost everything
    (c0, w0), (c1, w1) = orc.area_weights(1e-9, 8)
    assert (c0, c1) == (0, 1) and w1 == pytest.approx(1.0, abs=1e-7)


def test_area_weights_partition_of_unity(orc):
    rng = np.random.default_rng(1)
    for z in rng.random(200):
        for NZ in (8, 64, 256):
            (_, w0), (_, w1) = orc.area_weights(z, NZ)
            assert 0.0 <= w0 <= 1.0 and 0.0 <= w1 <= 1.0
            assert abs(w0 + w1 - 1.0) <= 2.3e-16


def test_gather_delta_and_constant_fields(orc):
    NX, NY = 8, 16
    rng = np.random.default_rng(2)
    for x, y in rng.random((50, 2)):
        (i0, a0), (i1, a1) = orc.area_weights(x, NX)
        (j0, b0), (j1, b1) = orc.area_weights(y, NY)
        # a unit field in one cell returns that cell's weight product exactly
        for (i, a) in ((i0, a0), (i1, a1)):
            for (j, b) in ((j0, b0), (j1, b1)):
                Ex = np.zeros((NY, NX))
                Ex[j, i] = 1.0
                ex, ey = orc.gather2d(Ex, -Ex, x, y)
                assert ex == a * b and ey == -(a * b)
        ex, ey = orc.gather2d(np.full((NY, NX), 3.0), np.full((NY, NX), -0.5), x, y)
        assert ex == pytest.approx(3.0, rel=1e-15) and ey == pytest.approx(-0.5, rel=1e-15)


def test_gather_is_a_fused_fold_in_reference_order(orc):
    # j outer, i inner, F = fma(E, wx*wy, F) (@muladd): checked against exact rationals
    NX = NY = 8
    rng = np.random.default_rng(3)
    Ex = rng.standard_normal((NY, NX))
    Ey = rng.standard_normal((NY, NX))
    for x, y in rng.random((100, 2)):
        (i0, a0), (i1, a1) = orc.area_weights(x, NX)
        (j0, b0), (j1, b1) = orc.area_weights(y, NY)
        acc = Fraction(0)
        accf = 0.0
        for (j, b) in ((j0, b0), (j1, b1)):
            for (i, a) in ((i0, a0), (i1, a1)):
                wxy = a * b
                acc = Fraction(accf) + Fraction(Ex[j, i]) * Fraction(wxy)   # one rounding per fma
                accf = float(acc)
        ex, _ = orc.gather2d(Ex, Ey, x, y)
        assert ex == accf


def textbook_boris(v, E, B, dt):
    """The textbook vector form (Birdsall & Langdon 4-3): half kick, rotation
    with t = B dt/2 and s = 2t/(1+t.t), half kick; q/m = 1."""
    v = np.array(v, dtype=np.longdouble)
    E = np.array(E, dtype=np.longdouble)
    t = np.array(B, dtype=np.longdouble) * dt / 2
    s = 2 * t / (1 + t @ t)
    vm = v + E * dt / 2
    vp = vm + np.cross(vm + np.cross(vm, t), s)
    return vp + E * dt / 2


def test_boris_against_textbook_vector_form(orc):
    rng = np.random.default_rng(4)
    for _ in range(200):
        v = rng.standard_normal(3)
        E2 = rng.standard_normal(2)
        B0, dt = rng.random() * 4, rng.random() * 0.2
        got = np.array(orc.boris(*v, *E2, dt, B0))
        ref = textbook_boris(v, [E2[0], E2[1], 0.0], [B0, 0.0, 0.0], dt).astype(np.float64)
        assert np.allclose(got, ref, rtol=0, atol=4e-16 * (1 + np.abs(ref).max()))


def test_boris_pure_rotation_about_x(orc):
    B0, dt = np.pi / 2, 1 / 1024
    vx, vy, vz = 0.7, 0.3, -0.4
    a, b, c = orc.boris(vx, vy, vz, 0.0, 0.0, dt, B0)
    assert a == vx                                    # B along x leaves vx alone
    assert np.hypot(b, c) == pytest.approx(np.hypot(vy, vz), rel=1e-15)
    ang = np.arctan2(c, b) - np.arctan2(vz, vy)
    assert ang == pytest.approx(-2 * np.arctan(B0 * dt / 2), rel=1e-12)   # v x B with B = +x


def test_deposit_exact_rational_sums(orc):
    NX, NY = 8, 4
    rng = np.random.default_rng(5)
    P = 300
    x, y = rng.random(P), rng.random(P)
    w = 0.7310585786300049
    rho = orc.deposit2d(NX, NY, x, y, w)
    exact = [[Fraction(0)] * NX for _ in range(NY)]
    for p in range(P):
        (i0, a0), (i1, a1) = orc.area_weights(x[p], NX)
        (j0, b0), (j1, b1) = orc.area_weights(y[p], NY)
        for (j, b) in ((j0, b0), (j1, b1)):
            for (i, a) in ((i0, a0), (i1, a1)):
                exact[j][i] += Fraction(a * b * w)           # each product rounds as in the script
    for j in range(NY):
        for i in range(NX):
            assert rho[j, i] == float(exact[j][i])
    assert rho.sum() == pytest.approx(P * w, rel=1e-14)
    # order- and thread-count independence
    assert np.array_equal(rho, orc.deposit2d(NX, NY, x, y, w, nthreads=4))
    perm = rng.permutation(P)
    assert np.array_equal(rho, orc.deposit2d(NX, NY, x[perm], y[perm], w))


def test_fft2d_against_numpy_and_bruteforce(orc):
    rng = np.random.default_rng(6)
    for NX, NY in ((4, 8), (16, 16), (64, 32)):
        re, im = rng.standard_normal((NY, NX)), rng.standard_normal((NY, NX))
        fr, fi = orc.fft2d(re, im)
        ref = np.fft.fft2(re + 1j * im)
        assert np.allclose(fr + 1j * fi, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        br, bi = orc.fft2d(re, im, inverse=1)
        ref = np.fft.ifft2(re + 1j * im) * NX * NY
        assert np.allclose(br + 1j * bi, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    # O(N^4) long-double DFT on a small grid
    NX, NY = 8, 4
    re, im = rng.standard_normal((NY, NX)), rng.standard_normal((NY, NX))
    fr, fi = orc.fft2d(re, im)
    jj, ii = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    z = (re + 1j * im).astype(np.clongdouble)
    out = np.zeros((NY, NX), dtype=np.clongdouble)
    for b in range(NY):
        for a in range(NX):
            ph = np.exp(-2j * np.pi * (np.longdouble(a) * ii / NX + np.longdouble(b) * jj / NY))
            out[b, a] = (z * ph).sum()
    assert np.allclose(fr + 1j * fi, out.astype(np.complex128), rtol=0, atol=1e-14 * np.abs(out).max())


@pytest.mark.parametrize("a,b", [(1, 0), (0, 1), (2, 3), (-3, 1), (5, -7)])
def test_fieldsolve_single_mode_closed_form(orc, a, b):
    # rho = cos(2 pi (a i/NX + b j/NY)) -> E = (kx, ky) sin(theta) / (kx^2 + ky^2)
    NX, NY = 32, 16
    jj, ii = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    th = 2 * np.pi * (a * ii / NX + b * jj / NY)
    kx, ky = 2 * np.pi * a, 2 * np.pi * b
    k2 = kx * kx + ky * ky
    Ex, Ey, phi = orc.fieldsolve2d(np.cos(th), want_phi=True)
    assert np.allclose(Ex, kx * np.sin(th) / k2, rtol=0, atol=1e-14)
    assert np.allclose(Ey, ky * np.sin(th) / k2, rtol=0, atol=1e-14)
    assert np.allclose(phi, np.cos(th), rtol=0, atol=1e-14)   # phis stores rho without its mean


def test_fieldsolve_mean_and_gauss_law(orc):
    NX, NY = 16, 32
    rng = np.random.default_rng(7)
    rho = 5.0 + rng.standard_normal((NY, NX))
    Ex, Ey, phi = orc.fieldsolve2d(rho, want_phi=True)
    assert abs(Ex.mean()) < 1e-15 and abs(Ey.mean()) < 1e-15
    assert np.allclose(phi, rho - rho.mean(), rtol=0, atol=1e-13)
    # spectral divergence of E equals rho - mean except on the Nyquist lines
    kx = 2 * np.pi * np.fft.fftfreq(NX, 1 / NX)
    ky = 2 * np.pi * np.fft.fftfreq(NY, 1 / NY)
    div = np.fft.ifft2(1j * kx[None, :] * np.fft.fft2(Ex) + 1j * ky[:, None] * np.fft.fft2(Ey))
    R = np.fft.fft2(rho - rho.mean())
    R[:, NX // 2] = 0
    R[NY // 2, :] = 0
    D = np.fft.fft2(div.real)
    D[:, NX // 2] = 0
    D[NY // 2, :] = 0
    assert np.allclose(D, R, rtol=0, atol=1e-11)


def test_diagnostics_hand_state(orc):
    NX, NY = 4, 2
    Ex = np.full((NY, NX), 3.0)
    Ey = np.full((NY, NX), 4.0)
    vx = np.array([1.0, -1.0, 2.0, 0.5])
    vy = np.array([0.0, 1.0, -1.0, 0.5])
    K = orc.diagnostics2d(Ex, Ey, vx, vy, 0.25)
    assert K[0] == 25.0
    assert K[1] == (1 + 2 + 5 + 0.5) * 0.25
    assert K[2] == K[0] + K[1]
    assert K[3] == 2.5 / 4 and K[4] == 0.5 / 4


def cold_lattice(NX, NY, ppc_side, eps, axis):
    n = NX * ppc_side
    m = NY * ppc_side
    x0 = (np.arange(n) + 0.5) / n
    y0 = (np.arange(m) + 0.5) / m
    X, Y = np.meshgrid(x0, y0, indexing="xy")
    X, Y = X.ravel(), Y.ravel()
    if axis == 0:
        X = X + eps * np.sin(2 * np.pi * X)
    else:
        Y = Y + eps * np.sin(2 * np.pi * Y)
    return X, Y


def oscillation_omega(K1, dt):
    """Angular frequency of a field oscillating as cos(w t): K1 ~ cos^2 peaks every pi/w."""
    k = np.arange(1, len(K1) - 1)
    pk = k[(K1[k] > K1[k - 1]) & (K1[k] >= K1[k + 1])]
    # parabolic refinement of each peak
    tp = []
    for i in pk:
        a, b, c = K1[i - 1], K1[i], K1[i + 1]
        tp.append(i + 0.5 * (a - c) / (a - 2 * b + c))
    tp = np.array(tp)
    return np.pi / (np.mean(np.diff(tp)) * dt)


@pytest.mark.parametrize("axis,name", [(0, "plasma"), (1, "upper_hybrid")])
def test_cold_plasma_frequencies(orc, axis, name):
    # k along x (parallel to B = B0 x): omega = omega_p = sqrt(n0);
    # k along y (perpendicular): omega = sqrt(n0 + B0^2).  Run with the textbook
    # pairing: the loop (push, deposit, solve, diagnostics) is the same code.
    NX = NY = 64
    n0 = 4 * np.pi ** 2
    B0 = np.sqrt(n0) / 4
    x, y = cold_lattice(NX, NY, 2, 1e-4, axis)
    P = x.size
    w = n0 / P * NX * NY
    dt = 1 / 400
    z = np.zeros(P)
    out = orc.run2d(NX, NY, dt, B0, w, x, y, z, z, z, 600, ns=1, nthreads=8, textbook=1)
    K = out[7]
    om = oscillation_omega(K[:, 0], dt)
    expect = np.sqrt(n0) if axis == 0 else np.sqrt(n0 + B0 ** 2)
    assert om == pytest.approx(expect, rel=0.005)
    # momentum along B is conserved to roundoff
    assert np.max(np.abs(K[:, 3])) < 1e-16


def test_reference_pairing_inverts_a_cold_lattice(orc):
    # With the reference's i / i+1 pairing the deposited charge centroid moves
    # against the particle inside a cell, so a cold lattice (no particle crosses
    # a cell edge) sees an inverted restoring force and the mode grows instead
    # of oscillating.  Pins the pairing's sign in a whole run.
    NX = NY = 64
    n0 = 4 * np.pi ** 2
    x, y = cold_lattice(NX, NY, 2, 1e-4, 0)
    P = x.size
    z = np.zeros(P)
    K = orc.run2d(NX, NY, 1 / 400, np.sqrt(n0) / 4, n0 / P * NX * NY, x, y, z, z, z, 200, ns=1,
                  nthreads=8)[7]
    assert K[-1, 0] > 50 * K[0, 0]
    assert np.all(np.diff(K[:, 0]) > 0)


def test_run_momentum_x_thermal(orc):
    NX = NY = 16
    rng = np.random.default_rng(8)
    P = NX * NY * 16
    x, y = rng.random(P), rng.random(P)
    vx = np.where(np.arange(P) < P // 2, -1.0, 1.0) + 0.1 * rng.standard_normal(P)
    vy, vz = 0.1 * rng.standard_normal(P), 0.1 * rng.standard_normal(P)
    n0 = 4 * np.pi ** 2
    out = orc.run2d(NX, NY, 1 / 64, np.sqrt(n0) / 4, n0 / P * NX * NY, x, y, vx, vy, vz, 100, ns=2,
                    nthreads=8)
    K = out[7]
    p0 = vx.sum() / P
    assert np.max(np.abs(K[:, 3] - p0)) < 1e-14
    assert np.all(np.isfinite(K))
    assert np.all(out[0] > 0) and np.all(out[0] <= 1) and np.all(out[1] > 0) and np.all(out[1] <= 1)
    # serial and threaded runs are bitwise equal (exact per-cell sums)
    out1 = orc.run2d(NX, NY, 1 / 64, np.sqrt(n0) / 4, n0 / P * NX * NY, x, y, vx, vy, vz, 20, ns=2)
    out8 = orc.run2d(NX, NY, 1 / 64, np.sqrt(n0) / 4, n0 / P * NX * NY, x, y, vx, vy, vz, 20, ns=2,
                     nthreads=8)
    for a, b in zip(out1, out8):
        assert np.array_equal(a, b)
