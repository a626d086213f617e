"""Pins the oracle of the 1D electromagnetic fixed-point script
(oracle/em1d.c, src/GaussianEMFixedPoint.jl) against what the mathematics and
the script's expressions fix: hand cases of the wrapped stencil centre,
scipy's erf, exact rational deposits, single-mode closed forms of the Gauss,
Ampere and Faraday updates, the velocity update's closed form, the vacuum
light wave's per-mode amplification matrix, the script's `Fz .*= Ez`
(Ez restarts each step), gyration in the background Bz, moments equal to
deposits of the final state, the stable sort, and split calls."""
from fractions import Fraction

import numpy as np
import pytest
from scipy.special import erf


def test_irn_hand_cases(orc):
    N = 64
    assert orc.em_irn(0.3, N) == 19                     # round(19.2)
    assert orc.em_irn(0.001, N) == N                    # round(0.064) = 0 -> mod1 -> N
    assert orc.em_irn(0.5 / N, N) == N                  # round(0.5) = 0 (ties to even)
    assert orc.em_irn(1.5 / N, N) == 2                  # round(1.5) = 2
    assert orc.em_irn(1.0, N) == N
    assert orc.em_irn(1.0 + 1.0 / N, N) == 1            # round(65) -> mod1 -> 1
    assert orc.em_irn(-0.2 / N, N) == N                 # round(-0.2) = -0.0 -> N


def test_stencil_weights_and_wrapped_centre(orc):
    N = 64
    rng = np.random.default_rng(0)
    for c in rng.uniform(1.0 / N, 1 - 1.0 / N, 100):
        cells, ff = orc.em_stencil(N, c)
        i0 = int(np.rint(c * N))
        i = np.arange(i0 - 6, i0 + 7)
        ref = (erf(((i + 0.5) / N - c) * N) - erf(((i - 0.5) / N - c) * N)) / 2
        assert np.allclose(ff, ref, rtol=0, atol=2e-16)
        assert np.array_equal(cells, (i - 1) % N)
        assert abs(ff.sum() - 1.0) < 1e-15
    # the centre is wrapped but c is not: no weight at all near x = 0 and past 1
    for c in (0.001, 0.4 / N, 1.0 + 0.7 / N):
        _, ff = orc.em_stencil(N, c)
        assert not np.any(ff)


def test_deposit_exact_rational_sums(orc):
    N = 32
    rng = np.random.default_rng(1)
    y = rng.uniform(0.05, 0.95, 60)
    u = rng.standard_normal(60)
    w = 3.25
    for uu in (None, u):
        rho = orc.em_deposit(N, y, w, uu)
        acc = [Fraction(0)] * N
        for j in range(60):
            cells, ff = orc.em_stencil(N, y[j])
            for k in range(13):
                a = (ff[k] if uu is None else uu[j] * ff[k]) * w
                acc[cells[k]] += Fraction(a)
        assert np.array_equal(rho, np.array([float(a) for a in acc]))
        tot = w * (60 if uu is None else u.sum())
        assert rho.sum() == pytest.approx(tot, rel=1e-12, abs=1e-12)


def test_fields_single_mode_closed_forms(orc):
    N, eps, dt = 32, 0.1, 1e-3
    j = np.arange(N)
    z = np.zeros(N)
    for m in (1, 3, 7):
        k = 2 * np.pi * m
        th = k * j / N
        # Gauss: rho = cos -> Ex = sin/k
        Ex, Ey, Ez, By, Bz, ey, ez, by, bz = orc.em_fields(eps, dt, np.cos(th), z, z, z, z, z, z, z, z)
        assert np.allclose(Ex, np.sin(th) / k, atol=1e-14)
        assert not np.any(Ey) and not np.any(By)
        # Ampere: bz = cos (Bz = Cz), no current -> Ey = Fy + dt*k*sin/eps^2
        c = np.cos(th)
        Ex, Ey, Ez, By, Bz, ey, ez, by, bz = orc.em_fields(eps, dt, z, z, z, z, z, z, c, z, c)
        assert np.allclose(Ey, dt * k * np.sin(th) / eps ** 2, atol=1e-12)
        # ... and the current enters with a minus sign
        Ex, Ey, *_ = orc.em_fields(eps, dt, z, c, z, z, z, z, z, z, z)
        assert np.allclose(Ey, -dt * c, atol=1e-15)
        # Faraday: ez = Fz = cos (Ez = Fz when by = Jz = 0) -> By = Cy - dt*k*sin
        Ex, Ey, Ez, By, Bz, ey, ez, by, bz = orc.em_fields(eps, dt, z, z, z, z, c, z, z, z, z)
        assert np.array_equal(Ez, c) and np.allclose(By, -dt * k * np.sin(th), atol=1e-14)
        assert np.allclose(by, By / 2, atol=0)
        # and ey = Fy = cos -> Bz = Cz + dt*k*sin
        Ex, Ey, Ez, By, Bz, *_ = orc.em_fields(eps, dt, z, z, z, c, z, z, z, z, z)
        assert np.allclose(Bz, dt * k * np.sin(th), atol=1e-14)


def test_push_closed_form(orc):
    N, dt = 32, 0.01
    rng = np.random.default_rng(2)
    for _ in range(20):
        E, B = rng.standard_normal(3), rng.standard_normal(2)
        c = rng.uniform(0.1, 0.9)
        u = rng.standard_normal(3)
        v = orc.em_push(c, *(np.full(N, e) for e in E), *(np.full(N, b) for b in B), dt, u)
        s = orc.em_stencil(N, c)[1].sum()
        a = E * s
        b = np.array([0.0, B[0] * s, B[1] * s])
        ref = u + (a + np.cross(u + a * dt / 2, b)) * dt
        assert np.allclose(v, ref, rtol=0, atol=1e-14)


def mode_amplitude(f, m):
    return np.fft.fft(f)[m] / len(f)


def test_vacuum_light_wave_amplification(orc):
    """No charge (w = 0): Ex stays zero, so every step is one sweep, and each
    Fourier mode of (Ey, Bz) evolves by the matrix the script's one pass gives:
    e' = e - i kap b/eps^2, b' = b - i kap (e + e')/2, kap = k dt."""
    N, eps, dt, T = 32, 0.1, 2e-4, 40
    j = np.arange(N)
    for m in (1, 2, 5):
        k = 2 * np.pi * m
        Ey0 = np.cos(k * j / N)
        E = np.vstack([np.zeros(N), Ey0, np.zeros(N)])
        B = np.vstack([np.zeros(N), 0.3 * np.sin(k * j / N)])
        x = np.array([0.5])
        v = np.zeros((3, 1))
        _, _, E1, B1, _, _, sweeps = orc.em_run(N, eps, dt, 0.0, x, v, T, E=E, B=B, want_fields=False)
        assert np.all(sweeps == 1)
        kap = k * dt
        M = np.array([[1, -1j * kap / eps ** 2], [-1j * kap, 1 - kap ** 2 / (2 * eps ** 2)]])
        s0 = np.array([mode_amplitude(Ey0, m), mode_amplitude(B[1], m)])
        s1 = np.linalg.matrix_power(M, T) @ s0
        assert np.allclose([mode_amplitude(E1[1], m), mode_amplitude(B1[1], m)], s1, rtol=1e-10, atol=1e-13)
        # |lambda|^2 = det M = 1 + kap^2/(2 eps^2) > 1: the single pass gains energy
        assert abs(np.linalg.det(M)) > 1


def test_ez_restarts_every_step(orc):
    """`Fz .*= Ez` where the script means `Fz .= Ez`: Fz stays zero, so a
    z-polarised wave with no current and no By is gone after one step."""
    N = 32
    j = np.arange(N)
    E = np.vstack([np.zeros(N), np.zeros(N), np.cos(2 * np.pi * j / N)])
    B = np.vstack([np.zeros(N), np.zeros(N)])
    _, _, E1, B1, Fz, _, _ = orc.em_run(N, 0.1, 1e-3, 0.0, np.array([0.5]), np.zeros((3, 1)), 1, E=E, B=B,
                                        want_fields=False)
    assert not np.any(E1[2]) and not np.any(B1[0]) and not np.any(Fz)


def test_gyration_in_the_background_field(orc):
    """w = 0 and the script's Bz = 1: v' = u + (u x b) dt with b = (0, 0, sum ff),
    one sweep per step: vx + i vy turns by (1 - i b dt) per step (and grows by
    sqrt(1 + (b dt)^2): the update is not a rotation)."""
    N, dt, T = 32, 0.01, 25
    rng = np.random.default_rng(3)
    x = rng.uniform(0.2, 0.8, 40)
    v = np.vstack([rng.standard_normal(40), rng.standard_normal(40), rng.standard_normal(40)])
    # positions move: keep them in the interior of the box over the run
    v[0] *= 0.01
    x1, v1, _, _, _, _, sweeps = orc.em_run(N, 0.1, dt, 0.0, x, v, T, sort=0, want_fields=False)
    assert np.all(sweeps == 1)
    z = v[0] + 1j * v[1]
    zz = z.copy()
    xx = x.copy()
    for _ in range(T):
        bz = np.array([orc.em_stencil(N, xx[p] + zz[p].real * dt / 2)[1].sum() for p in range(40)])
        xx = xx + zz.real * dt
        zz = zz * (1 - 1j * bz * dt)
    assert np.allclose(v1[0] + 1j * v1[1], zz, rtol=1e-12, atol=1e-14)
    assert np.array_equal(v1[2], v[2])


def script_setup(N=32, P=256, seed=4):
    eps = 0.1
    dt = eps ** 2 / N / 2
    W = 400
    w = W / P * N
    rng = np.random.default_rng(seed)
    x = rng.random(P)
    a = 2 * np.pi * rng.random(P)
    v = np.vstack([np.sin(a) / 4, np.cos(a) / 4, np.zeros(P)])
    return N, eps, dt, w, x, v


def test_moments_sort_and_split_calls(orc):
    N, eps, dt, w, x, v = script_setup()
    T = 6
    x1, v1, E1, B1, Fz1, f, sweeps = orc.em_run(N, eps, dt, w, x, v, T)
    assert np.all((sweeps >= 2) & (sweeps <= 3))
    assert np.all(np.diff(x1) >= 0)
    # the moment fields of the last step are deposits of the final state
    assert np.array_equal(f[5, -1], orc.em_deposit(N, x1, w))
    for k in range(3):
        assert np.array_equal(f[6 + k, -1], orc.em_deposit(N, x1, w, v1[k]))
        assert np.array_equal(f[9 + k, -1], orc.em_deposit(N, x1, w, v1[k] * v1[k]))
    assert np.array_equal(f[0, -1], E1[0]) and np.array_equal(f[4, -1], B1[1])
    # two calls equal one
    a = orc.em_run(N, eps, dt, w, x, v, 2)
    b = orc.em_run(N, eps, dt, w, a[0], a[1], T - 2, E=a[2], B=a[3], Fz=a[4])
    for p, q in zip(b[:5], (x1, v1, E1, B1, Fz1)):
        assert np.array_equal(p, q)
    assert np.array_equal(np.concatenate([a[5], b[5]], axis=1), f)


def test_sort_is_stable_argsort(orc):
    rng = np.random.default_rng(5)
    x = rng.choice([0.1, 0.2, 0.2, 0.7, 0.0, -0.0, 1.0], 500)
    v = rng.standard_normal((3, 500))
    xs, vs, perm = orc.em_sort(x, v)
    order = np.argsort(x, kind="stable")
    # numpy orders -0.0 and 0.0 as equal; isless puts -0.0 first
    keyed = sorted(range(500), key=lambda j: (x[j], not np.signbit(x[j]), j))
    assert np.array_equal(perm, keyed)
    assert np.array_equal(xs, x[keyed]) and np.array_equal(vs, v[:, keyed])
    assert np.array_equal(np.abs(xs), np.abs(x[order]))
