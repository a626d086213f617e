"""Mixed precision (FP32 push and shape weights, exact integer charge sums,
FP64 field solve / convergence test / diagnostics) against the FP64 path.

Bars (the accuracy the mixed form is held to): field-energy relative error <= 1e-6 over
the run and the fitted growth rate within 1% of FP64; results reproducible
bit for bit from run to run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def run(pg, cfg, x, v, nsteps):
    with pg.Context(cfg) as c:
        c.upload(x, v)
        D, S = c.run(nsteps)
        xo, vo, Eo = c.download()
    return D[0], S[0], xo, vo, Eo


def test_upload_download_round_trip(pg):
    from picgolf import configs
    cfg, x, v, _ = configs.scaled_two_stream(precision="mixed", P=4096)
    with pg.Context(cfg) as c:
        c.upload(x, v)
        xo, vo, _ = c.download()
    assert np.max(np.abs(xo - x)) <= 2.0 ** -33          # 2^-32 fixed point, rounded
    assert np.max(np.abs(vo - v) / np.abs(v)) <= 2.0 ** -47       # hi + lo carries 48 bits


def test_mixed_is_deterministic(pg):
    from picgolf import configs
    cfg, x, v, _ = configs.scaled_two_stream(precision="mixed", P=2 ** 20)
    a = run(pg, cfg, x, v, 20)
    b = run(pg, cfg, x, v, 20)
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


def test_mixed_field_energy_and_momentum_vs_fp64_full_size(pg):
    """The scaled two-stream workload at full size: N=4096, P=2^24, 300 steps."""
    from picgolf import configs
    n = 300
    cm, x, v, _ = configs.scaled_two_stream(precision="mixed")
    cf, _, _, _ = configs.scaled_two_stream(precision="fp64")
    Dm, Sm, *_ = run(pg, cm, x, v, n)
    Df, Sf, *_ = run(pg, cf, x, v, n)
    rel = np.max(np.abs(Dm[:, 0] - Df[:, 0]) / Df[:, 0])
    print(f"field-energy rel err {rel:.3e}; sweeps mixed {Sm.mean():.2f} fp64 {Sf.mean():.2f}")
    assert rel <= 1e-6
    assert np.max(np.abs(Dm[:, 3] - Dm[0, 3])) < 1e-9      # momentum drift
    assert abs(Sm.mean() - Sf.mean()) <= 0.5


def test_mixed_quiet_two_stream_growth_rate(pg):
    """Quiet start (N=64, P=2048).  Held on exact fixed-point lattices the
    mixed path's quiet start stays exactly uniform (E = 0), so both runs get
    the same seed: a 1e-8 mode-1 displacement.  Both must grow at the analytic
    two-stream rate and agree with each other to 1%."""
    from picgolf import configs
    cm, x, v, _ = configs.quiet_two_stream(precision="mixed")
    cm.rtol = 1e-8
    cf, _, _, _ = configs.quiet_two_stream(precision="fp64")
    cf.rtol = 1e-8
    x = np.mod(x + 1e-8 * np.sin(2 * np.pi * x), 1.0)
    n = 3000
    Dm, *_ = run(pg, cm, x, v, n)
    Df, *_ = run(pg, cf, x, v, n)
    t = np.arange(1, n + 1) * cf.dt
    W = cf.W
    xk = 2 * np.pi / np.sqrt(W / 2)
    gamma = np.imag(np.sqrt(complex(xk ** 2 + 1 - np.sqrt(4 * xk ** 2 + 1)))) * np.sqrt(W / 2)
    slope_pred = 2 * gamma / np.log(10)
    slopes = []
    for D in (Dm, Df):
        la = np.log10(D[:, 0])
        sel = (la > la[0] + 2) & (la < -3)
        assert sel.sum() > 200
        slopes.append(np.polyfit(t[sel], la[sel], 1)[0])
    print("growth slopes (decades/time): mixed %.4f fp64 %.4f analytic %.4f" % (*slopes, slope_pred))
    assert slopes[0] == pytest.approx(slopes[1], rel=0.01)
    assert slopes[1] == pytest.approx(slope_pred, rel=0.02)


def test_mixed_speculative_slots_match_full_slots():
    """The mixed path through the halt-and-finish route equals a context that
    runs every sweep slot (bit for bit)."""
    import picgolf
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_two_stream(precision="mixed")
    N = cfg.N
    j = np.arange(N)
    E0 = 40.0 * np.sin(2 * np.pi * j / N)
    rng = np.random.default_rng(12)
    v2 = v + 0.5 * rng.standard_normal(v.size)
    outs = []
    for warm in (False, True):
        with picgolf.Context(cfg) as c:
            if warm:
                c.upload(x, v)
                c.run(20)                   # learns a small slot count
            c.upload(x, v2, E0)
            D, S = c.run(12)
            outs.append((D, S, c.download()))
    (Da, Sa, A), (Db, Sb, B) = outs
    assert np.array_equal(Sa, Sb) and np.array_equal(Da, Db)
    for a, b in zip(A, B):
        assert np.array_equal(a, b)


def test_fused_first_deposit_matches_a_fresh_start():
    """The final sweep deposits the next step's first charge; a context that
    starts from the downloaded state (and so deposits it in a launch of its
    own) continues bit for bit."""
    import picgolf
    from picgolf import configs
    cfg, x, v, _ = configs.scaled_two_stream(precision="mixed", P=2 ** 18, N=1024)
    with picgolf.Context(cfg) as c:
        c.upload(x, v)
        D, S = c.run(20)
        A = c.download()
    with picgolf.Context(cfg) as c:
        c.upload(x, v)
        D1, S1 = c.run(10)
        xm, vm, Em = c.download()
    with picgolf.Context(cfg) as c:
        c.upload(xm, vm, Em)
        D2, S2 = c.run(10)
        B = c.download()
    assert np.array_equal(S[0], np.concatenate([S1[0], S2[0]]))
    assert np.array_equal(D[0, :, 0], np.concatenate([D1[0, :, 0], D2[0, :, 0]]))
    for a, b in zip(A, B):
        assert np.array_equal(a, b)
