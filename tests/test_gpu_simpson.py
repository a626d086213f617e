"""GPU parity of the Simpson-1/3 fixed point (src/GaussianFixedPointQuietSimpson13.jl,
src/AreaFixedPointQuietSimpson13.jl) through the C ABI: whole FP64 runs bit
for bit against the CPU restatement for the Gaussian and area shapes, batches
of uneven instances, and the scripts' quiet-start growth on the device."""
import numpy as np
import pytest

from test_oracle_physics import gamma_decades

pytestmark = pytest.mark.gpu

SHAPE_CODE = {"gauss7": 7, "gauss6": 6, "area": 1}


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("shape", ["gauss7", "area", "gauss6"])
def test_quiet_simpson_run_bitwise(pg, orc, shape):
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_simpson("area" if shape == "area" else "gauss7")
    cfg.shape = shape
    T1, T2 = 23, 37                     # two calls: graphs of every slot, then of the sweeps seen
    with pg.Context(cfg) as c:
        c.upload(x, v)
        D1, S1 = c.run(T1)
        D2, S2 = c.run(T2)
        xg, vg, Eg = c.download()
    xo, vo, E3, Do, So = orc.run_simpson(cfg.N, SHAPE_CODE[shape], cfg.W, cfg.w, cfg.dt, x, v, T1 + T2,
                                         cfg.rtol, nthreads=8)
    assert np.array_equal(np.concatenate([S1[0], S2[0]]), So)
    assert same(np.vstack([D1[0], D2[0]]), Do)
    assert same(xg, xo) and same(vg, vo) and same(Eg, E3[2])
    assert So.min() >= 2


def test_simpson_batch_equals_independent_runs(pg, orc):
    from picgolf.core import Config
    cfgs, states = [], []
    for k, (N, ppc, W, shape) in enumerate([(32, 20, 150.0, "area"), (256, 16, 300.0, "area"),
                                            (64, 33, 220.0, "area"), (1024, 8, 400.0, "area")]):
        P = N * ppc
        cfgs.append(Config(N=N, P=P, W=W, dt=1 / (6 * N), shape=shape, scheme="simpson", rtol=1e-12))
        r = np.random.default_rng(k)
        states.append((r.random(P), np.where(np.arange(P) < P // 2, -1.0, 1.0) + 0.05 * r.standard_normal(P)))
    T = 12
    with pg.Context(cfgs) as c:
        for i, (x, v) in enumerate(states):
            c.upload(x, v, inst=i)
        D, S = c.run(T)
        for i, cfg in enumerate(cfgs):
            xg, vg, Eg = c.download(i)
            xo, vo, E3, Do, So = orc.run_simpson(cfg.N, 1, cfg.W, cfg.w, cfg.dt, *states[i], T, cfg.rtol,
                                                 nthreads=8)
            assert np.array_equal(S[i], So), i
            for a, b in zip((xg, vg, Eg, D[i]), (xo, vo, E3[2], Do)):
                assert same(a, b), i


def test_simpson_tiling_and_autotune_leave_results_unchanged(pg):
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_simpson("gauss7")
    outs = []
    for mode in ("default", "small", "tuned"):
        with pg.Context(cfg) as c:
            c.upload(x, v)
            if mode == "small":
                c.set_tiling(64, 64)
            elif mode == "tuned":
                c.autotune(trial_steps=2)
            D, S = c.run(20)
            outs.append((D, S, c.download()))
    for D, S, st in outs[1:]:
        assert np.array_equal(D, outs[0][0]) and np.array_equal(S, outs[0][1])
        for a, b in zip(st, outs[0][2]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", ["gauss7", "area"])
def test_quiet_simpson_growth_rate_on_device(pg, shape):
    # the figure of the scripts: ~30 decades of growth on the analytic 2 gamma slope
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_simpson(shape)
    nsteps = 2600
    with pg.Context(cfg) as c:
        c.upload(x, v)
        D, S = c.run(nsteps)
    D = D[0]
    t = np.arange(1, nsteps + 1) * cfg.dt
    la = np.log10(np.abs(D[:, 0]))
    sel = (t > 1.5) & (t < 6.5)
    slope = np.polyfit(t[sel], la[sel], 1)[0]
    print(f"Simpson {shape}: growth slope {slope:.4f} decades/time (analytic {2 * gamma_decades(cfg.W):.4f}), "
          f"mean sweeps {S.mean():.2f}")
    assert slope == pytest.approx(2 * gamma_decades(cfg.W), rel=0.01)
    assert np.max(np.abs(D[:, 3])) < 1e-14


@pytest.mark.parametrize("shape", ["area", "gauss7", "gauss6"])
def test_simpson_operators_bitwise(pg, orc, shape):
    """The script's closures one at a time: rho(a, b) with the area or Gaussian
    shape and the three stage kicks of a sweep."""
    from picgolf.core import Config
    code = SHAPE_CODE[shape]
    N, P = 64, 5000
    rng = np.random.default_rng(code)
    X = rng.random(P)
    x = X + 0.02 * rng.standard_normal(P)          # unwrapped, as inside the sweeps
    V, v = rng.standard_normal(P), rng.standard_normal(P)
    E = rng.standard_normal(N)
    w, dt = 0.37, 1 / 384
    with pg.Context(Config(N=N, P=P, W=100.0, dt=dt, shape=shape, scheme="simpson")) as c:
        for a, b in ((X, X), (X, x), (x, x)):
            assert same(c.deposit(a, b, w=w, shape=shape), orc.deposit(N, code, a, b, w=w))
        for a, b, vin, coef in ((X, X, V, dt), (X, x, v, 4 * dt), (x, x, v, dt)):
            got = c.simpson_kick(a, b, vin, E, coef)
            ref = np.array([vin[j] + (orc.gather_at(N, code, E, (a[j] + b[j]) / 2) * coef) / 6 for j in range(P)])
            assert same(got, ref)
