"""GPU parity of the magnetised 1D2V path (src/NGP1D2V.jl, src/NGP1D2V2S.jl)
through the C ABI in include/pic1d2v.h: whole runs of both scripts' set-ups,
the binned field and the diagnostics bit for bit against the CPU restatement,
the operators one at a time, and the upper-hybrid oscillation on the device."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def oracle_run(orc, cfg, st, T):
    xs, vxs, vys = [s[0] for s in st], [s[1] for s in st], [s[2] for s in st]
    return orc.run_1d2v(cfg.N, 7, cfg.q, cfg.q_m, cfg.mass, cfg.w, cfg.n0, cfg.B0, cfg.dt, cfg.bin, xs, vxs, vys, T,
                        nthreads=16)


@pytest.mark.parametrize("species,T1,T2", [(1, 17, 23), (2, 45, 30)])
def test_bernstein_run_bitwise(pg, orc, species, T1, T2):
    from picgolf import configs
    cfg, st, _ = configs.bernstein(species)
    with pg.Context1D2V(cfg) as c:
        for s, (x, vx, vy) in enumerate(st):
            c.upload(s, x, vx, vy)
        D = np.vstack([c.run_raw(T1), c.run_raw(T2)])
        Es = c.fields_raw()
        got = [c.download(s) for s in range(len(st))]
    xs, vxs, vys, Eo, Do = oracle_run(orc, cfg, st, T1 + T2)
    assert D.shape[0] == (T1 + T2) // cfg.bin
    assert same(D, Do[:D.shape[0]])
    assert same(Es, Eo)
    for s in range(len(st)):
        assert same(got[s][0], xs[s]) and same(got[s][1], vxs[s]) and same(got[s][2], vys[s])


def test_operators_bitwise(pg, orc):
    from picgolf import configs
    cfg, st, _ = configs.bernstein(2)
    rng = np.random.default_rng(5)
    with pg.Context1D2V(cfg) as c:
        xs = [s[0] for s in st]
        rho = c.deposit(xs)
        assert same(rho, orc.deposit_1d2v(cfg.N, 7, cfg.q, cfg.w, xs))
        E = orc.fieldsolve(rho)
        x = rng.uniform(-0.01, 1.01, 3000)
        Ej = c.gather(E, x)
        assert same(Ej, [orc.gather_at(cfg.N, 7, E, xj) for xj in x])
        vx, vy = rng.standard_normal(3000), rng.standard_normal(3000)
        for qm in (-1.0, 1 / 8):
            a, b = vx.copy(), vy.copy()
            c.boris(a, b, Ej, q_m=qm)
            ref = np.array([orc.boris12(vx[j], vy[j], Ej[j], cfg.B0, cfg.dt, qm) for j in range(3000)])
            assert same(a, ref[:, 0]) and same(b, ref[:, 1])


def test_upload_restarts_the_run(pg):
    from picgolf import configs
    cfg, st, _ = configs.bernstein(1)
    with pg.Context1D2V(cfg) as c:
        c.upload(0, *st[0])
        D1 = c.run_raw(40)
        E1 = c.fields_raw()
        c.upload(0, *st[0])                       # back to t = 0: bins and rows start again
        D2 = c.run_raw(40)
        E2 = c.fields_raw()
    assert np.array_equal(D1, D2) and np.array_equal(E1, E2)
    with pg.Context1D2V(cfg) as c:
        with pytest.raises(pg.PicError) as e:
            c.run_raw(1)
        assert e.value.code == -5


def test_upper_hybrid_oscillation_on_device(pg):
    from picgolf.core1d2v import Config1D2V
    N = 64
    P = 32 * N
    n0 = 4 * np.pi ** 2
    B0 = 2 * np.pi
    dt = 1 / (6 * N)
    cfg = Config1D2V(N=N, P=P, bin=1, w=n0 / P * N, n0=n0, B0=B0, dt=dt)
    x0 = (np.arange(P) + 0.5) / P
    z = np.zeros(P)
    T = 1200
    with pg.Context1D2V(cfg) as c:
        c.upload(0, x0 + 1e-4 * np.sin(2 * np.pi * x0), z, z)
        D = c.run_raw(T)
    d1 = D[:, 0]
    mins = [i for i in range(1, T - 1) if d1[i] < d1[i - 1] and d1[i] <= d1[i + 1]]
    omega = 2 * np.pi / (np.mean(np.diff(mins)) * dt)
    k = 2 * np.pi
    S = np.exp(-k ** 2 / (4 * N ** 2)) * np.sin(k / (2 * N)) / (k / (2 * N))
    assert omega == pytest.approx(np.sqrt(n0 * S ** 2 + B0 ** 2), rel=3e-3)
