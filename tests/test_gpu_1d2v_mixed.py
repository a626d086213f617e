"""The mixed-precision form of the magnetised 1D2V scripts (src/NGP1D2V.jl,
src/NGP1D2V2S.jl) against their FP64 path over the scripts' whole runs
(2^14 and 2^16 steps), with the errors stated.

The FP32 force (weights and field rounded to FP32, about 6e-8 of a kick each
step) accumulates over these long runs: the field-energy error stays near 1e-6
for the first eighth of each run and reaches ~1e-5 (one species) and ~2e-4 (two
species) at the end.  FP64 runs from positions or velocities perturbed by 1e-9
drift apart by at most 2e-7 over the same runs, so the drift is the precision's
and not chaos.  The binned field and its omega-k power, which the scripts plot,
agree to 4e-6 / 2e-7 (one species) and 1.3e-4 / 5e-5 (two species)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def full_run(pg, species, precision):
    from picgolf import configs
    cfg, st, T = configs.bernstein(species)
    cfg.precision = precision
    with pg.Context1D2V(cfg) as c:
        for s, (x, vx, vy) in enumerate(st):
            c.upload(s, x, vx, vy)
        D = c.run_raw(T)
        Es = c.fields_raw()
        got = [c.download(s) for s in range(len(st))]
    return D, Es, got


# (species, field-energy bar over the first eighth, over the run, Es, omega-k power)
BARS = [(1, 2e-6, 2e-5, 1e-5, 1e-6), (2, 2e-6, 5e-4, 4e-4, 2e-4)]


@pytest.mark.parametrize("species,b8,ball,bes,bpow", BARS)
def test_mixed_bernstein_full_run_against_fp64(pg, species, b8, ball, bes, bpow):
    Df, Ef, _ = full_run(pg, species, "fp64")
    Dm, Em, gm = full_run(pg, species, "mixed")
    r1 = np.abs(Dm[:, 0] - Df[:, 0]) / np.abs(Df[:, 0])
    r2 = np.abs(Dm[:, 1] - Df[:, 1]) / np.abs(Df[:, 1])
    es = np.linalg.norm(Em - Ef) / np.linalg.norm(Ef)
    Pf, Pm = np.abs(np.fft.fft2(Ef)) ** 2, np.abs(np.fft.fft2(Em)) ** 2
    pw = np.linalg.norm(Pm - Pf) / np.linalg.norm(Pf)
    dmom = np.abs(Dm[:, 3:5] - Df[:, 3:5]).max()
    n8 = len(r1) // 8
    print(f"1D2V species {species} mixed vs FP64: field energy {r1[:n8].max():.3e} (first eighth), "
          f"{r1.max():.3e} (run); kinetic {r2.max():.3e}; Es {es:.3e}; omega-k power {pw:.3e}; "
          f"momentum rows {dmom:.2e} (FP64 drift {np.abs(Df[:, 3] - Df[0, 3]).max():.2e})")
    assert r1[:n8].max() <= b8 and r1.max() <= ball
    assert r2.max() <= 1e-6
    assert es <= bes and pw <= bpow
    assert dmom <= 1e-3 * np.abs(Df[:, 3:5]).max()
    for x, vx, vy in gm:
        assert np.all((x >= 0) & (x < 1)) and np.all(np.isfinite(vx)) and np.all(np.isfinite(vy))


def test_mixed_upper_hybrid_and_determinism(pg):
    from picgolf.core1d2v import Config1D2V
    N = 64
    P = 32 * N
    n0 = 4 * np.pi ** 2
    B0 = 2 * np.pi
    dt = 1 / (6 * N)
    cfg = Config1D2V(N=N, P=P, bin=1, w=n0 / P * N, n0=n0, B0=B0, dt=dt, precision="mixed")
    x0 = (np.arange(P) + 0.5) / P
    z = np.zeros(P)
    T = 1200
    runs = []
    for _ in range(2):
        with pg.Context1D2V(cfg) as c:
            c.upload(0, x0 + 1e-4 * np.sin(2 * np.pi * x0), z, z)
            runs.append((c.run_raw(T), c.download(0)))
    assert np.array_equal(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)
    d1 = runs[0][0][:, 0]
    mins = [i for i in range(1, T - 1) if d1[i] < d1[i - 1] and d1[i] <= d1[i + 1]]
    omega = 2 * np.pi / (np.mean(np.diff(mins)) * dt)
    k = 2 * np.pi
    S = np.exp(-k ** 2 / (4 * N ** 2)) * np.sin(k / (2 * N)) / (k / (2 * N))
    assert omega == pytest.approx(np.sqrt(n0 * S ** 2 + B0 ** 2), rel=3e-3)


def test_mixed_upload_download_round_trip(pg):
    from picgolf import configs
    cfg, st, _ = configs.bernstein(2)
    cfg.precision = "mixed"
    with pg.Context1D2V(cfg) as c:
        for s, (x, vx, vy) in enumerate(st):
            c.upload(s, x, vx, vy)
        for s, (x, vx, vy) in enumerate(st):
            gx, gvx, gvy = c.download(s)
            assert np.max(np.abs(gx - x)) <= 2.0 ** -33
            assert np.max(np.abs(gvx - vx)) <= 2.0 ** -46 * np.max(np.abs(vx))
            assert np.max(np.abs(gvy - vy)) <= 2.0 ** -46 * np.max(np.abs(vy))
