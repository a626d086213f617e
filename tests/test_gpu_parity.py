"""GPU parity: every FP64 operator and whole runs through the C ABI, compared
bit for bit with the CPU restatement (the oracle) on the same inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float64).eps)


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def ctx_for(pg, N, shape="gauss7", scheme="midpoint", P=64, W=100.0):
    from picgolf.core import Config
    return pg.Context(Config(N=N, P=P, W=W, dt=1 / (6 * N), shape=shape, scheme=scheme))


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("N,R,P", [(16, 7, 1), (16, 7, 37), (64, 7, 5000), (64, 6, 5000),
                                   (4096, 7, 100000), (32, 6, 20000)])
def test_deposit_gauss_bitwise(pg, orc, N, R, P):
    rng = np.random.default_rng(N * 7 + P)
    x = rng.uniform(-0.05, 1.05, P)          # unwrapped midpoints reach outside [0, 1)
    X = rng.random(P)
    w = 400.0 / P * N
    with ctx_for(pg, N) as c:
        for y in (X, None):
            got = c.deposit(x, y, w=w, shape=R)
            ref = orc.deposit(N, R, x, y, w=w)
            assert same(got, ref)


def test_deposit_ngp_bitwise_and_ties(pg, orc):
    N, P = 128, 8192
    rng = np.random.default_rng(0)
    x = rng.random(P)
    x[:64] = (np.arange(64) + 0.5) / N        # exact ties: round half to even
    x[64] = 0.0
    x[65] = 1.0
    with ctx_for(pg, N, shape="ngp", scheme="leapfrog") as c:
        got = c.deposit(x, w=3.125, shape="ngp")
    assert same(got, orc.deposit(N, 0, x, w=3.125))


@pytest.mark.parametrize("N", [8, 16, 32, 64, 128, 256, 1024, 2048, 4096])
def test_fieldsolve_bitwise(pg, orc, N):
    rng = np.random.default_rng(N)
    rho = 100 + rng.standard_normal(N)
    with ctx_for(pg, N) as c:
        assert same(c.fieldsolve(rho), orc.fieldsolve(rho))
        j = np.arange(N)
        assert same(c.fieldsolve(np.cos(2 * np.pi * j / N)), orc.fieldsolve(np.cos(2 * np.pi * j / N)))


@pytest.mark.parametrize("R", [6, 7])
def test_push_midpoint_and_position_bitwise(pg, orc, R):
    N, P = 64, 20000
    rng = np.random.default_rng(R)
    X = rng.random(P)
    V = np.where(np.arange(P) < P // 2, -1.0, 1.0) + 0.1 * rng.standard_normal(P)
    v = V + 1e-3 * rng.standard_normal(P)
    dt = 1 / (6 * N)
    E = rng.standard_normal(N)
    with ctx_for(pg, N) as c:
        x = c.midpoint_position(X, V, v, dt)
        assert same(x, orc.midpoint_position(X, V, v, dt))
        vv = np.zeros(P)
        c.push_midpoint(X, V, x, vv, E, dt, shape=R)
        assert same(vv, orc.push_midpoint(N, R, X, V, x, E, dt))


@pytest.mark.parametrize("shape,R", [("ngp", 0), ("gauss6", 6)])
def test_leapfrog_pieces_bitwise(pg, orc, shape, R):
    N, P = 128, 8192
    rng = np.random.default_rng(3)
    x = rng.random(P)
    x[:10] = -1e-20                             # mod(-tiny, 1) == 1.0
    v = np.where(np.arange(P) < P // 2, -1.0, 1.0)
    E = rng.standard_normal(N)
    dt = 1 / (4 * N)
    with ctx_for(pg, N, shape=shape, scheme="leapfrog") as c:
        xd = x.copy()
        c.drift_half(xd, v, dt)
        assert same(xd, orc.drift_half(x, v, dt))
        xg, vg = xd.copy(), v.copy()
        c.push_leapfrog(xg, vg, E, dt, shape=R)
        xo = orc.drift_half(xd, v, dt)
        assert same(xg, xo)
        assert same(vg, orc.kick(N, R, v, xo, E, dt))
        xw = np.array([-1e-20, 1.0, 1.5, -0.25, 0.3, 7.75, -3.5])
        c.wrap(xw)
        assert same(xw, orc.wrap(np.array([-1e-20, 1.0, 1.5, -0.25, 0.3, 7.75, -3.5])))


def test_isapprox_and_diagnostics_bitwise(pg, orc):
    N = 64
    rng = np.random.default_rng(1)
    E = rng.standard_normal(N) * 1e-17
    with ctx_for(pg, N) as c:
        cases = [(np.full(N, np.nan), E, 1e-8), (E, E, 0.0), (E * (1 + 1e-9), E, 1e-8),
                 (E * (1 + 1e-9), E, 1e-10), (np.zeros(N), E, 4 * EPS), (E, E + 1e-33, 4 * EPS),
                 (np.full(N, 1e-200), np.full(N, 1e-200) * (1 + 1e-12), 1e-8)]
        for F, G, rtol in cases:
            assert c.isapprox(F, G, rtol) == orc.isapprox(F, G, rtol)
        for P in (2048, 4096, 4097, 300000):
            v = rng.standard_normal(P)
            assert same(c.diagnostics(E, v, 100.0), orc.diagnostics(E, v, 100.0))


def run_both(pg, orc, cfg, x, v, nsteps):
    with pg.Context(cfg) as c:
        c.upload(x, v)
        D, S = c.run(nsteps)
        xg, vg, Eg = c.download()
    shape = {"ngp": 0, "gauss6": 6, "gauss7": 7}[cfg.shape]
    scheme = 1 if cfg.scheme == "midpoint" else 0
    xo, vo, Eo, Do, So = orc.run(cfg.N, shape, scheme, cfg.W, cfg.w, cfg.dt, x, v, nsteps,
                                 rtol=cfg.rtol, max_sweeps=cfg.max_sweeps, nthreads=8)
    return (xg, vg, Eg, D[0], S[0]), (xo, vo, Eo, Do, So)


def test_run_quiet_two_stream_bitwise(pg, orc):
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_two_stream()
    g, o = run_both(pg, orc, cfg, x, v, 300)
    assert np.array_equal(g[4], o[4])            # same sweep counts
    for a, b in zip(g[:4], o[:4]):
        assert same(a, b)


def test_run_ngp_two_stream_bitwise(pg, orc):
    from picgolf import configs
    cfg, x, v, _ = configs.ngp_two_stream()
    g, o = run_both(pg, orc, cfg, x, v, 200)
    for a, b in zip(g[:4], o[:4]):
        assert same(a, b)


def test_run_gaussian_leapfrog_bitwise(pg, orc):
    # src/Gaussian.jl: NX=128, NP=64NX, dt=1/10NX, W=1600, w=W/NP, deposit k[2]*w/dx, +-6
    from picgolf.core import Config
    N, P, W = 128, 64 * 128, 1600.0
    cfg = Config(N=N, P=P, W=W, w=W / P * N, dt=1 / (10 * N), shape="gauss6", scheme="leapfrog")
    rng = np.random.default_rng(2)
    x = rng.random(P)
    v = np.where(np.arange(1, P + 1) > P / 2, 1.0, -1.0)
    g, o = run_both(pg, orc, cfg, x, v, 40)
    for a, b in zip(g[:4], o[:4]):
        assert same(a, b)


def test_run_scaled_config_reduced_particles_bitwise(pg, orc):
    from picgolf import configs
    cfg, x, v, _ = configs.scaled_two_stream(precision="fp64", P=2 ** 15)
    g, o = run_both(pg, orc, cfg, x, v, 3)
    assert np.array_equal(g[4], o[4])
    for a, b in zip(g[:4], o[:4]):
        assert same(a, b)


def test_batch_equals_independent_runs(pg, orc):
    from picgolf.core import Config
    cfgs, states = [], []
    for k, (N, ppc, W) in enumerate([(32, 20, 150.0), (128, 64, 300.0), (64, 33, 220.0)]):
        P = N * ppc
        cfgs.append(Config(N=N, P=P, W=W, dt=1 / (6 * N), rtol=1e-8))
        r = np.random.default_rng(k)
        states.append((r.random(P), np.where(np.arange(P) < P // 2, -1.0, 1.0)))
    with pg.Context(cfgs) as c:
        for i, (x, v) in enumerate(states):
            c.upload(x, v, inst=i)
        D, S = c.run(25)
        for i, cfg in enumerate(cfgs):
            xg, vg, Eg = c.download(i)
            xo, vo, Eo, Do, So = orc.run(cfg.N, 7, 1, cfg.W, cfg.w, cfg.dt, *states[i], 25, rtol=cfg.rtol, nthreads=8)
            assert np.array_equal(S[i], So)
            for a, b in zip((xg, vg, Eg, D[i]), (xo, vo, Eo, Do)):
                assert same(a, b)


def test_full_size_scaled_fp64_properties(pg):
    """The scaled two-stream size (N=4096, P=2^24): momentum conserved, run-to-run bitwise
    reproducible (no float atomics anywhere)."""
    from picgolf import configs
    cfg, x, v, _ = configs.scaled_two_stream(precision="fp64")
    outs = []
    for _ in range(2):
        with pg.Context(cfg) as c:
            c.upload(x, v)
            D, S = c.run(2)
            outs.append((D, c.download()))
    assert same(outs[0][0], outs[1][0]) and same(outs[0][1][0], outs[1][1][0])
    D = outs[0][0][0]
    assert np.all(np.isfinite(D))
    p0 = np.sum(v / cfg.P)
    assert abs(D[-1, 3] - p0) < 1e-13


def test_speculative_sweep_slots_recover_bitwise(pg, orc):
    """After a first run the step graphs hold only as many sweep slots as the
    loop has needed so far.  A later state that needs more sweeps halts the
    graph and finishes the step on the host side: results stay bit-identical
    to the oracle and the sweep counts are the oracle's."""
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_two_stream()
    N = cfg.N
    with pg.Context(cfg) as c:
        c.upload(x, v)
        _, S1 = c.run(20)
        # a strong field and hot particles: more Picard sweeps per step
        j = np.arange(N)
        E0 = 40.0 * np.sin(2 * np.pi * j / N) + 15.0 * np.cos(6 * np.pi * j / N)
        rng = np.random.default_rng(11)
        v2 = v + 0.5 * rng.standard_normal(v.size)
        c.upload(x, v2, E0)
        D, S2 = c.run(12)
        xg, vg, Eg = c.download()
    xo, vo, Eo, Do, So = orc.run(N, 7, 1, cfg.W, cfg.w, cfg.dt, x, v2, 12, rtol=cfg.rtol, E=E0)
    assert S2.max() > S1.max()                      # the halt-and-finish path ran
    assert np.array_equal(S2[0], So)
    assert same(D[0], Do) and same(xg, xo) and same(vg, vo) and same(Eg, Eo)
