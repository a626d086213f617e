"""Pins the Simpson-1/3 fixed-point restatement (oracle/simpson.c) and the
linear area weights of src/AreaFixedPointQuietSimpson13.jl:5 against hand
cases, a transcription of the script's loop (src/GaussianFixedPointQuietSimpson13.jl:8-18)
over the separately pinned primitives, the analytic two-stream growth rate the
script overplots (:26-27) and momentum conservation."""
import numpy as np
import pytest

from test_oracle_physics import EPS, gamma_decades

AREA = 1


def test_area_stencil_hand_cases(orc):
    N = 8
    # y*N = 3.25: grid points 4/N (weight 0.25) and 3/N (0.75); Julia index i is 0-based i-1
    assert orc.area_stencil(N, 3.25 / 8) == ((3, 0.25), (2, 0.75))
    assert orc.area_stencil(N, 4 / 8) == ((3, 1.0), (2, 0.0))
    # y = 0 and y = 1 are the grid point N/N: mod1(0, N) = N
    assert orc.area_stencil(N, 0.0) == ((7, 1.0), (6, 0.0))
    assert orc.area_stencil(N, 1.0) == ((7, 1.0), (6, 0.0))
    # an unwrapped position just below 0 (x is wrapped only after the sweeps)
    (c0, w0), (c1, w1) = orc.area_stencil(N, -2.0 ** -12)
    assert (c0, c1) == (7, 6) and w1 == 2.0 ** -9 and w0 == 1 - 2.0 ** -9
    # and just above 1
    (c0, w0), (c1, w1) = orc.area_stencil(N, 1 + 2.0 ** -12)
    assert (c0, c1) == (0, 7) and w0 == 2.0 ** -9 and w1 == 1 - 2.0 ** -9


def test_area_weights_interpolate_linear_fields(orc):
    N = 16
    rng = np.random.default_rng(0)
    a = 2.5
    E = a * (np.arange(N) + 1.0) / N            # cell i (0-based) sits at (i+1)/N
    for c in rng.uniform(1.0 / N, 1.0, 300):      # away from the wrap of the linear ramp
        assert orc.gather_at(N, AREA, E, c) == pytest.approx(a * c, rel=1e-14, abs=1e-15)
        assert orc.gather_at(N, AREA, np.ones(N), c) == pytest.approx(1.0, abs=2 * EPS)


def test_gauss_gather_is_the_midpoint_kick(orc):
    N, P = 32, 500
    rng = np.random.default_rng(1)
    E = rng.standard_normal(N)
    X = rng.random(P)
    dt = 1 / 192
    v = orc.push_midpoint(N, 7, X, np.zeros(P), X, E, dt)   # V + s*dt with V = 0, centre (X+X)/2
    g = np.array([orc.gather_at(N, 7, E, (X[j] + X[j]) / 2) for j in range(P)])
    assert np.array_equal(v, g * dt)


def transcribed_simpson(orc, N, shape, W, w, dt, x, v, T, rtol, max_sweeps=10):
    """src/GaussianFixedPointQuietSimpson13.jl:8-18, line by line, over the
    oracle's primitives (exact deposit, field solve, isapprox over the 3 x N
    matrix in column-major order, scalar gathers)."""
    P = x.size
    E = np.zeros((3, N))
    D, S = [], []
    for _ in range(T):
        X, V = x.copy(), v.copy()
        F = np.full((3, N), np.nan)
        E[0] = orc.fieldsolve(orc.deposit(N, shape, X, X, w))
        n = 0
        while n < max_sweeps:
            if orc.isapprox(F.T.ravel(), E.T.ravel(), rtol):
                break
            F = E.copy()
            x = orc.midpoint_position(X, V, v, dt)
            v = np.array([V[j] + (orc.gather_at(N, shape, E[0], (X[j] + X[j]) / 2) * dt) / 6 for j in range(P)])
            E[1] = orc.fieldsolve(orc.deposit(N, shape, X, x, w))
            v = np.array([v[j] + (orc.gather_at(N, shape, E[1], (X[j] + x[j]) / 2) * (4 * dt)) / 6
                          for j in range(P)])
            E[2] = orc.fieldsolve(orc.deposit(N, shape, x, x, w))
            v = np.array([v[j] + (orc.gather_at(N, shape, E[2], (x[j] + x[j]) / 2) * dt) / 6 for j in range(P)])
            n += 1
        x = orc.wrap(x)
        D.append(orc.diagnostics(E[2], v, W))
        S.append(n)
    return x, v, E, np.array(D), np.array(S)


@pytest.mark.parametrize("shape,rtol", [(7, 4 * EPS), (AREA, 1e-14), (6, 1e-12)])
def test_simpson_run_matches_the_transcribed_loop(orc, shape, rtol):
    N, P, T = 16, 96, 5
    W = 32 * np.pi ** 2 / 3
    w = W / P * N
    dt = 1 / (6 * N)
    x, v = orc.quiet_start(P)
    x = x + 0.01 * np.sin(2 * np.pi * x)          # a seeded perturbation, so the field is not zero
    x = x - np.floor(x)
    ref = transcribed_simpson(orc, N, shape, W, w, dt, x, v, T, rtol)
    for nthreads in (1, 4):
        got = orc.run_simpson(N, shape, W, w, dt, x, v, T, rtol, nthreads=nthreads)
        for g, r in zip(got, ref):
            assert np.array_equal(np.asarray(g), np.asarray(r))
    assert ref[4].min() >= 2


@pytest.mark.parametrize("shape,rtol", [(7, 4 * EPS), (AREA, 1e-14)])
def test_simpson_quiet_start_growth_rate_and_momentum(orc, shape, rtol):
    # the scripts' configuration: N=64, P=32N, dt=1/6N, W=32pi^2/3
    N = 64
    P = 32 * N
    dt = 1 / (6 * N)
    W = 32 * np.pi ** 2 / 3
    w = W / P * N
    x, v = orc.quiet_start(P)
    nsteps = 2600
    _, _, _, D, sw = orc.run_simpson(N, shape, W, w, dt, x, v, nsteps, rtol, nthreads=8)
    t = np.arange(1, nsteps + 1) * dt
    la = np.log10(np.abs(D[:, 0]))
    assert la[0] < -28                            # roundoff-seeded start
    sel = (t > 1.5) & (t < 6.5)
    slope = np.polyfit(t[sel], la[sel], 1)[0]
    assert slope == pytest.approx(2 * gamma_decades(W), rel=0.01)
    assert np.max(np.abs(D[:, 3])) < 1e-14        # each stage deposits and gathers at one centre
    assert 1 <= sw.min() and sw.max() <= 10
