"""Run-level physics checks that pin the oracle's whole step loop (A9-A15)
against what the mathematics fixes: the analytic two-stream growth rate the
reference overplots (src/GaussianFixedPointQuiet.jl:19-20), momentum
conservation of the midpoint scheme (README.md:45-47) and the cold plasma
frequency sqrt(W)."""
import numpy as np
import pytest

EPS = 2.220446049250313e-16


def gamma_decades(W, m=1, u=1.0):
    """gamma(x) of src/GaussianFixedPointQuiet.jl:19 with x = 2 pi m u / sqrt(W/2)."""
    x = 2 * np.pi * m * u / np.sqrt(W / 2)
    return np.imag(np.sqrt(complex(x ** 2 + 1 - np.sqrt(4 * x ** 2 + 1)))) * np.sqrt(W / 2) / np.log(10)


def test_analytic_gamma_closed_form():
    W = 32 * np.pi ** 2 / 3
    assert gamma_decades(W) * np.log(10) == pytest.approx(2 * np.pi / np.sqrt(3), rel=1e-14)


def test_quiet_start_growth_rate_and_momentum(orc):
    # config: N=64, P=32N, dt=1/6N, W=32pi^2/3, +-7 stencil, rtol=4eps
    N = 64
    P = 32 * N
    dt = 1 / (6 * N)
    W = 32 * np.pi ** 2 / 3
    w = W / P * N
    x, v = orc.quiet_start(P)
    nsteps = 2600
    _, _, _, D, sw = orc.run(N, 7, 1, W, w, dt, x, v, nsteps, rtol=4 * EPS, nthreads=8)
    t = np.arange(1, nsteps + 1) * dt
    la = np.log10(np.abs(D[:, 0]))
    assert -36 < la[0] < -31                      # roundoff-seeded start, ~1e-33 (figure)
    sel = (t > 1.5) & (t < 6.5)
    slope = np.polyfit(t[sel], la[sel], 1)[0]
    assert slope == pytest.approx(2 * gamma_decades(W), rel=0.01)
    assert np.max(np.abs(D[:, 3])) < 1e-14         # momentum ~1e-16 throughout
    assert 1 <= sw.min() and sw.max() <= 10 and sw.mean() > 2


def test_cold_plasma_oscillation_frequency(orc):
    N = 64
    P = 32 * N
    dt = 1 / (6 * N)
    W = 32 * np.pi ** 2 / 3
    w = W / P * N
    j = np.arange(P)
    x0 = (j + 0.5) / P
    x = x0 + 1e-4 * np.sin(2 * np.pi * x0)
    v = np.zeros(P)
    nsteps = 1000
    _, _, _, D, _ = orc.run(N, 7, 1, W, w, dt, x, v, nsteps, rtol=1e-13, nthreads=8)
    d1 = D[:, 0]
    mins = [i for i in range(1, nsteps - 1) if d1[i] < d1[i - 1] and d1[i] <= d1[i + 1]]
    # field energy ~ cos^2(omega t): minima every pi/omega
    half_period = np.mean(np.diff(mins)) * dt
    omega = np.pi / half_period
    # cell-integrated Gaussian of width 1/(sqrt2 N) deposited and gathered: S(k)^2
    k = 2 * np.pi
    S = np.exp(-k ** 2 / (4 * N ** 2)) * np.sin(k / (2 * N)) / (k / (2 * N))
    assert omega == pytest.approx(np.sqrt(W) * S, rel=3e-3)


def test_ngp_leapfrog_runs_and_free_streams(orc):
    N = 128
    P = 64 * N
    dt = 1 / (4 * N)
    W = 200.0
    w = W / P * N
    rng = np.random.default_rng(0)
    x = rng.random(P)
    v = np.where(np.arange(P) < P // 2, -1.0, 1.0)
    x1, v1, E, D, _ = orc.run(N, 0, 0, W, w, dt, x, v, 64)
    assert np.all((x1 >= 0) & (x1 <= 1))
    assert np.all(np.isfinite(D)) and D[0, 1] == pytest.approx(1.0, rel=1e-3)
    # one particle per cell centre, all at rest: rho is uniform, E = 0, nothing moves
    xs = (np.arange(N) + 1.0) / N
    xs[-1] = 0.0
    x2, v2, E2, _, _ = orc.run(N, 0, 0, W, W, dt, xs, np.zeros(N), 8)
    assert np.array_equal(x2, xs) and np.max(np.abs(v2)) < 1e-15
