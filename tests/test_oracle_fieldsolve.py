"""Pins the oracle's FFT and spectral Poisson solve (A5)."""
import numpy as np
import pytest


@pytest.mark.parametrize("N", [4, 8, 16, 64, 128])
def test_twiddles_accurate_and_symmetric(orc, N):
    W = orc.twiddles(N)
    m = np.arange(N // 2).astype(np.longdouble)
    pi = np.longdouble("3.14159265358979323846264338327950288")
    th = 2 * pi * m / N
    assert np.max(np.abs(W.real - np.cos(th))) < 1.2e-16
    assert np.max(np.abs(W.imag + np.sin(th))) < 1.2e-16
    if N >= 4:
        assert W[N // 4] == complex(0.0, -1.0)
    assert W[0] == complex(1.0, 0.0)


@pytest.mark.parametrize("N", [8, 32, 64, 256])
def test_fft_against_bruteforce_dft(orc, N):
    rng = np.random.default_rng(N)
    z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for inv in (False, True):
        got = orc.fft(z, inverse=inv)
        ref = orc.dft_bruteforce(z, inverse=inv)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("N", [64, 1024, 4096])
def test_fft_against_numpy(orc, N):
    rng = np.random.default_rng(N + 1)
    z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    assert np.allclose(orc.fft(z), np.fft.fft(z), rtol=0, atol=1e-12 * np.sqrt(N))
    assert np.allclose(orc.fft(z, inverse=True), np.fft.ifft(z) * N, rtol=0, atol=1e-12 * np.sqrt(N))


@pytest.mark.parametrize("N,m", [(64, 1), (64, 5), (128, 31), (4096, 2), (4096, 2047)])
def test_fieldsolve_single_mode_closed_form(orc, N, m):
    # rho_j = cos(2 pi m j / N) => E_j = sin(2 pi m j / N) / (2 pi m)
    j = np.arange(N)
    rho = np.cos(2 * np.pi * m * j / N)
    E = orc.fieldsolve(rho)
    ref = np.sin(2 * np.pi * m * j / N) / (2 * np.pi * m)
    assert np.max(np.abs(E - ref)) < 1e-15 * max(1.0, np.log2(N))


@pytest.mark.parametrize("N", [16, 64, 512])
def test_fieldsolve_zeroes_mean_and_nyquist(orc, N):
    j = np.arange(N)
    assert np.max(np.abs(orc.fieldsolve(np.full(N, 7.25)))) < 1e-15
    # the Nyquist term is purely imaginary after /ik, and real() drops it
    assert np.max(np.abs(orc.fieldsolve((-1.0) ** j))) < 1e-15


def test_fieldsolve_linear_and_matches_numpy_restatement(orc):
    N = 256
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal(N), rng.standard_normal(N)
    Ea, Eb = orc.fieldsolve(a), orc.fieldsolve(b)
    assert np.allclose(orc.fieldsolve(2 * a + 3 * b), 2 * Ea + 3 * Eb, atol=1e-14)
    # independent evaluation with numpy's FFT (different rounding)
    m = np.concatenate([[1], np.arange(1, N // 2 + 1), np.arange(-N // 2 + 1, 0)])
    xi = np.fft.fft(a) / (2j * np.pi * m)
    xi[0] = 0
    ref = np.real(np.fft.ifft(xi))
    assert np.max(np.abs(Ea - ref)) < 1e-15
    # Gauss's law: dE/dx = rho - mean(rho), spectrally
    k = 2 * np.pi * np.fft.fftfreq(N, 1 / N)
    dE = np.real(np.fft.ifft(1j * k * np.fft.fft(Ea)))
    nyq = np.real(np.fft.fft(a)[N // 2]) * (-1.0) ** np.arange(N) / N
    assert np.max(np.abs(dE - (a - a.mean() - nyq))) < 1e-12
