"""Pins the oracle of the 1D electromagnetic explicit script
(oracle/em1d.c orc_emx_*, src/GaussianEMExplicit.jl) against what the
mathematics and the script's expressions fix: single-mode closed forms of
Gauss and of both curls (numpy's FFT), the split kick's closed form in complex
notation, the vacuum leapfrog's per-mode matrix (determinant one: the explicit
scheme keeps the light wave's energy), a numpy restatement of whole steps that
tells the script's unsorted X in (X .+ x)/2 apart from a sorted one, moments
equal to deposits of the final state, and split calls."""
import numpy as np
import pytest


def modes(N):
    return np.concatenate([np.arange(0, N // 2 + 1), np.arange(-N // 2 + 1, 0)])


def test_gauss_and_faraday_single_modes(orc):
    N, dt = 32, 1e-3
    j = np.arange(N)
    z = np.zeros(N)
    for m in (1, 4, 9):
        k = 2 * np.pi * m
        th = k * j / N
        # Gauss: rho = cos -> Ex = sin/k; B untouched without Ey, Ez
        E, B, b2 = orc.emx_gauss_faraday(dt, np.cos(th), np.zeros((3, N)), np.vstack([z, np.ones(N)]))
        assert np.allclose(E[0], np.sin(th) / k, atol=1e-14)
        assert not np.any(B[0]) and np.array_equal(B[1], np.ones(N)) and np.array_equal(b2[1], np.ones(N))
        # By -= dt*rif(-ik.*fft(Ez)): Ez = cos -> By = -dt*k*sin
        E, B, b2 = orc.emx_gauss_faraday(dt, z, np.vstack([z, z, np.cos(th)]), np.zeros((2, N)))
        assert np.allclose(B[0], -dt * k * np.sin(th), atol=1e-14) and not np.any(B[1])
        assert np.array_equal(b2[0], B[0] / 2)
        # Bz -= dt*rif(ik.*fft(Ey)): Ey = cos -> Bz = +dt*k*sin
        E, B, b2 = orc.emx_gauss_faraday(dt, z, np.vstack([z, np.cos(th), z]), np.zeros((2, N)))
        assert np.allclose(B[1], dt * k * np.sin(th), atol=1e-14) and not np.any(B[0])
        # the old E is what Faraday reads: Ey, Ez are not changed
        assert np.array_equal(E[1], np.cos(th)) and not np.any(E[2])


def test_ampere_single_modes(orc):
    N, eps, dt = 32, 0.1, 1e-3
    j = np.arange(N)
    z = np.zeros(N)
    for m in (1, 4, 9):
        k = 2 * np.pi * m
        th = k * j / N
        c = np.cos(th)
        # Ey += dt*rif(-ik.*fft(Bz)/eps^2): Bz = cos -> +dt*k*sin/eps^2
        E = orc.emx_ampere(eps, dt, z, z, np.zeros((3, N)), np.vstack([z, c]))
        assert np.allclose(E[1], dt * k * np.sin(th) / eps ** 2, atol=1e-12) and not np.any(E[2])
        # Ez += dt*rif(+ik.*fft(By)/eps^2): By = cos -> -dt*k*sin/eps^2
        E = orc.emx_ampere(eps, dt, z, z, np.zeros((3, N)), np.vstack([c, z]))
        assert np.allclose(E[2], -dt * k * np.sin(th) / eps ** 2, atol=1e-12) and not np.any(E[1])
        # the currents enter with a minus sign, and Ex is left alone
        E0 = np.vstack([c, c, c])
        E = orc.emx_ampere(eps, dt, c, 2 * c, E0, np.zeros((2, N)))
        assert np.allclose(E[1], c - dt * c, atol=1e-15) and np.allclose(E[2], c - 2 * dt * c, atol=1e-15)
        assert np.array_equal(E[0], c)


def test_split_kick_closed_form(orc):
    """E kick a dt/2, then v += v x b dt, then a dt/2.  With b along z the
    middle step is (vx + i vy)(1 - i b dt); with b along y it is
    (vz + i vx)(1 - i b dt)."""
    N, dt = 32, 0.02
    rng = np.random.default_rng(2)
    for _ in range(20):
        Ev = rng.standard_normal(3)
        beta = rng.standard_normal()
        x = rng.uniform(0.1, 0.9)
        u = rng.standard_normal(3)
        S = orc.em_stencil(N, x)[1].sum()
        a = Ev * S
        E = np.vstack([np.full(N, e) for e in Ev])
        for axis in ("z", "y"):
            b2 = np.vstack([np.full(N, beta if axis == "y" else 0.0), np.full(N, beta if axis == "z" else 0.0)])
            v = orc.emx_push(x, E, b2, dt, u)
            h = u + a * dt / 2
            if axis == "z":
                zc = (h[0] + 1j * h[1]) * (1 - 1j * beta * S * dt)
                mid = np.array([zc.real, zc.imag, h[2]])
            else:
                zc = (h[2] + 1j * h[0]) * (1 - 1j * beta * S * dt)
                mid = np.array([zc.imag, h[1], zc.real])
            assert np.allclose(v, mid + a * dt / 2, rtol=0, atol=1e-14)


def mode_amplitude(f, m):
    return np.fft.fft(f)[m] / len(f)


@pytest.mark.parametrize("pol", ["z", "y"])
def test_vacuum_leapfrog_matrix(orc, pol):
    """No charge (w = 0): per Fourier mode, Faraday then Ampere is
    b' = b + s i kap e, e' = e + s i kap b'/eps^2 (s = +1 for (Ez, By),
    -1 for (Ey, Bz)), kap = k dt: det = 1, so for kap < 2 eps the wave keeps
    its amplitude."""
    N, eps, dt, T = 32, 0.1, 2e-4, 60
    j = np.arange(N)
    z = np.zeros(N)
    for m in (1, 3, 6):
        k = 2 * np.pi * m
        e0, b0 = np.cos(k * j / N), 0.3 * np.sin(k * j / N)
        if pol == "z":
            E, B, s = np.vstack([z, z, e0]), np.vstack([b0, z]), 1
        else:
            E, B, s = np.vstack([z, e0, z]), np.vstack([z, b0]), -1
        _, _, E1, B1, _ = orc.emx_run(N, eps, dt, 0.0, np.array([0.5]), np.zeros((3, 1)), T, E=E, B=B, sort=0,
                                      want_fields=False)
        kap = k * dt
        M = np.array([[1 - kap ** 2 / eps ** 2, s * 1j * kap / eps ** 2], [s * 1j * kap, 1]])
        assert abs(np.linalg.det(M) - 1) < 1e-12
        ie, ib = (2, 0) if pol == "z" else (1, 1)
        s0 = np.array([mode_amplitude(e0, m), mode_amplitude(b0, m)])
        s1 = np.linalg.matrix_power(M, T) @ s0
        got = np.array([mode_amplitude(E1[ie], m), mode_amplitude(B1[ib], m)])
        assert np.allclose(got, s1, rtol=1e-10, atol=1e-13)
        assert not np.any(E1[0]) or np.max(np.abs(E1[0])) == 0


def numpy_step(orc, N, eps, dt, w, x, v, E, B, sort_X):
    """One step of the explicit script in numpy (stencils from the pinned
    orc.em_stencil, sums in floating point, numpy's FFT).  sort_X: permute X
    with x, which the script does not do."""
    ik = 2j * np.pi * modes(N)

    def dep(y, u):
        r = np.zeros(N)
        for p in range(y.size):
            cells, ff = orc.em_stencil(N, y[p])
            np.add.at(r, cells, u[p] * ff * w)
        return r

    def rif(a):
        return np.fft.ifft(a).real

    X = x.copy()
    xi = np.fft.fft(dep(X, np.ones(x.size)))
    xi[1:] = xi[1:] / ik[1:]
    xi[0] = 0
    Ex = rif(xi)
    x = np.mod(X + v[0] * dt, 1.0)
    p = np.argsort(x, kind="stable")
    x, v = x[p], v[:, p]
    if sort_X:
        X = X[p]
    Ey, Ez = E[1], E[2]
    By = B[0] - dt * rif(-ik * np.fft.fft(Ez))
    Bz = B[1] - dt * rif(ik * np.fft.fft(Ey))
    by, bz = (By + B[0]) / 2, (Bz + B[1]) / 2
    v = v.copy()
    for q in range(x.size):
        cells, ff = orc.em_stencil(N, x[q])
        a = np.array([Ex[cells] @ ff, Ey[cells] @ ff, Ez[cells] @ ff])
        b = np.array([0.0, by[cells] @ ff, bz[cells] @ ff])
        u = v[:, q] + a * dt / 2
        u = u + np.cross(u, b) * dt
        v[:, q] = u + a * dt / 2
    c = (X + x) / 2
    Jy, Jz = dep(c, v[1]), dep(c, v[2])
    Ey = Ey + dt * rif(-ik * np.fft.fft(Bz) / eps ** 2 - np.fft.fft(Jy))
    Ez = Ez + dt * rif(ik * np.fft.fft(By) / eps ** 2 - np.fft.fft(Jz))
    return x, v, np.vstack([Ex, Ey, Ez]), np.vstack([By, Bz])


def script_setup(N=32, P=128, seed=6):
    eps = 0.1
    dt = eps ** 2 / N / 2
    W = 400
    w = W / P * N
    rng = np.random.default_rng(seed)
    return N, eps, dt, w, rng.random(P), rng.standard_normal((3, P))


def test_whole_steps_against_numpy_and_the_unsorted_X(orc):
    N, eps, dt, w, x, v = script_setup()
    T = 3
    xo, vo, Eo, Bo, _ = orc.emx_run(N, eps, dt, w, x, v, T, want_fields=False)
    res = {}
    for sort_X in (False, True):
        xx, vv, E, B = x.copy(), v.copy(), np.zeros((3, N)), np.vstack([np.zeros(N), np.ones(N)])
        for _ in range(T):
            xx, vv, E, B = numpy_step(orc, N, eps, dt, w, xx, vv, E, B, sort_X)
        res[sort_X] = (xx, vv, E, B)
    xx, vv, E, B = res[False]
    assert np.array_equal(xo, xx)
    assert np.allclose(vo, vv, rtol=1e-11, atol=1e-12)
    assert np.allclose(Eo, E, rtol=1e-10, atol=1e-9 * np.abs(E).max())
    assert np.allclose(Bo, B, rtol=1e-10, atol=1e-12)
    # a sorted X gives other currents: the oracle follows the script's pairing
    E_sorted = res[True][2]
    assert np.max(np.abs(E_sorted[1:] - Eo[1:])) > 1e3 * np.max(np.abs(E[1:] - Eo[1:])) + 1e-9


def test_moments_sort_and_split_calls(orc):
    N, eps, dt, w, x, v = script_setup(P=200, seed=7)
    T = 5
    x1, v1, E1, B1, f = orc.emx_run(N, eps, dt, w, x, v, T)
    assert np.all(np.diff(x1) >= 0)
    assert np.array_equal(f[5, -1], orc.em_deposit(N, x1, w))
    for k in range(3):
        assert np.array_equal(f[6 + k, -1], orc.em_deposit(N, x1, w, v1[k]))
        assert np.array_equal(f[9 + k, -1], orc.em_deposit(N, x1, w, v1[k] * v1[k]))
    assert np.array_equal(f[0, -1], E1[0]) and np.array_equal(f[2, -1], E1[2]) and np.array_equal(f[3, -1], B1[0])
    a = orc.emx_run(N, eps, dt, w, x, v, 2)
    b = orc.emx_run(N, eps, dt, w, a[0], a[1], T - 2, E=a[2], B=a[3])
    for p, q in zip(b[:4], (x1, v1, E1, B1)):
        assert np.array_equal(p, q)
    assert np.array_equal(np.concatenate([a[4], b[4]], axis=1), f)
    # without the sort, the particles keep their order and their positions
    # move by v1*dt exactly
    x2, v2, *_ = orc.emx_run(N, eps, dt, w, x, v, 1, sort=0, want_fields=False)
    assert np.array_equal(x2, np.mod(x + v[0] * dt, 1.0))
