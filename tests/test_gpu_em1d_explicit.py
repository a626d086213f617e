"""GPU parity of the 1D electromagnetic explicit script
(src/GaussianEMExplicit.jl) through the C ABI in include/picem1d.h (scheme
PICEM1D_EXPLICIT): the field updates (Gauss, Faraday, Ampere) and the split
kick bit for bit against the CPU restatement; whole runs of the script's
set-up (fields[:, t, 1:12] and the final state) over two calls, with and
without the per-step sort; a 2^14-particle run at N = 256; and, at N = 1024
with 2^20 particles, the size-independent properties that the moment fields
are deposits of the final state, the unsorted move is x = mod(x + v1*dt, 1)
bit for bit, and two calls equal one."""
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


def setup(N=64, P=1024, seed=0):
    """The explicit script's set-up: v = randn(P, 3), Bz = 1."""
    eps = 0.1
    dt = eps ** 2 / N / 2
    W = 400
    w = W / P * N
    rng = np.random.default_rng(seed)
    return N, eps, dt, w, rng.random(P), rng.standard_normal((3, P))


def cfg(N, P, eps, dt, w, sort=True):
    from picgolf.core_em1d import ConfigEM1D
    return ConfigEM1D(N=N, P=P, eps=eps, dt=dt, w=w, sort=sort, scheme="explicit")


def test_fields_and_push_bitwise(pg, orc):
    N, eps, dt, w, x, v = setup()
    rng = np.random.default_rng(2)
    rho, Jy, Jz = (rng.standard_normal(N) for _ in range(3))
    E, B = rng.standard_normal((3, N)), rng.standard_normal((2, N))
    with pg.ContextEM1D(cfg(N, 1024, eps, dt, w)) as c:
        E1, B1, b2 = c.explicit_fields(rho, Jy, Jz, E, B)
        Eo, Bo, bo = orc.emx_gauss_faraday(dt, rho, E, B)
        Eo = orc.emx_ampere(eps, dt, Jy, Jz, Eo, Bo)
        assert same(E1, Eo) and same(B1, Bo) and same(b2, bo)
        xs = rng.uniform(-0.01, 1.01, 2000)
        xs[:2] = [0.2 / N, 1.0 + 0.6 / N]          # wrapped-centre positions: no field
        u = rng.standard_normal((3, 2000))
        vv = c.explicit_push(xs, E1, b2, u)
        ref = np.array([orc.emx_push(xs[j], E1, b2, dt, u[:, j]) for j in range(2000)]).T
        assert same(vv, ref)
        assert same(vv[:, :2], u[:, :2] + 0.0)


@pytest.mark.parametrize("P", [1024, 8192])
@pytest.mark.parametrize("sort", [1, 0])
def test_run_bitwise_over_two_calls(pg, orc, sort, P):
    """The scripts' P = 1024, and 8192 (eight deposit blocks per sweep)."""
    N, eps, dt, w, x, v = setup(P=P)
    T1, T2 = 4, 6
    with pg.ContextEM1D(cfg(N, P, eps, dt, w, sort=bool(sort))) as c:
        c.upload(x, v)
        f1, s1 = c.run(T1)
        f2, s2 = c.run(T2)
        xg, vg, Eg, Bg = c.download()
    xo, vo, Eo, Bo, fo = orc.emx_run(N, eps, dt, w, x, v, T1 + T2, sort=sort)
    assert np.all(s1 == 1) and np.all(s2 == 1)
    assert same(np.concatenate([f1, f2], axis=1), fo)
    assert same(xg, xo) and same(vg, vo) and same(Eg, Eo) and same(Bg, Bo)


def test_run_mid_size_and_state_roundtrip(pg, orc):
    N, P = 256, 1 << 14
    _, eps, dt, w, x, v = setup(N=N, P=P, seed=3)
    rng = np.random.default_rng(3)
    E, B = 0.1 * rng.standard_normal((3, N)), np.vstack([0.1 * rng.standard_normal(N), np.ones(N)])
    with pg.ContextEM1D(cfg(N, P, eps, dt, w)) as c:
        c.upload(x, v, E, B)
        xg, vg, Eg, Bg = c.download()
        assert same(xg, x) and same(vg, v) and same(Eg, E) and same(Bg, B)
        f, s = c.run(3)
        xg, vg, Eg, Bg = c.download()
    xo, vo, Eo, Bo, fo = orc.emx_run(N, eps, dt, w, x, v, 3, E=E, B=B)
    assert same(f, fo)
    assert same(xg, xo) and same(vg, vo) and same(Eg, Eo) and same(Bg, Bo)


def test_properties_at_scale(pg, orc):
    """N = 1024, P = 2^20: the moment fields of the last step are the GPU
    deposits of the final state; without the sort one step moves x to
    mod(x + v1*dt, 1) exactly; two calls equal one."""
    N, P = 1024, 1 << 20
    _, eps, dt, w, x, v = setup(N=N, P=P, seed=4)
    with pg.ContextEM1D(cfg(N, P, eps, dt, w)) as c:
        c.upload(x, v)
        f, s = c.run(3)
        x1, v1, E1, B1 = c.download()
        assert np.all(np.diff(x1) >= 0) and np.all(s == 1)
        assert same(f[5, -1], c.deposit(x1, w))
        for k in range(3):
            assert same(f[6 + k, -1], c.deposit(x1, w, v1[k]))
            assert same(f[9 + k, -1], c.deposit(x1, w, v1[k] * v1[k]))
        assert same(f[0, -1], E1[0]) and same(f[1, -1], E1[1]) and same(f[4, -1], B1[1])
        c.upload(x, v)
        fa, _ = c.run(1)
        fb, _ = c.run(2)
        assert same(np.concatenate([fa, fb], axis=1), f)
        assert same(c.download()[0], x1)
    with pg.ContextEM1D(cfg(N, P, eps, dt, w, sort=False)) as c:
        c.upload(x, v)
        c.run(1, want_fields=False)
        xu = c.download()[0]
    assert same(xu, np.mod(x + v[0] * dt, 1.0))
