"""GPU parity of the 1D electromagnetic fixed-point path
(src/GaussianEMFixedPoint.jl) through the C ABI in include/picem1d.h: the
deposit (with the wrapped-centre particles), one sweep's field updates and the
velocity update bit for bit against the CPU restatement; whole runs of the
script's set-up (fields[:, t, 1:12], sweeps, the final state) over two calls,
with and without the per-step sort; and, at N = 1024 with 2^20 particles, the
size-independent property that the sort changes no field bit and only
permutes the particles."""
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
    eps = 0.1
    dt = eps ** 2 / N / 2
    W = 400
    w = W / P * N
    rng = np.random.default_rng(seed)
    x = rng.random(P)
    a = 2 * np.pi * rng.random(P)
    v = np.vstack([np.sin(a) / 4, np.cos(a) / 4, np.zeros(P)])
    return N, eps, dt, w, x, v


def test_deposit_bitwise(pg, orc):
    from picgolf.core_em1d import ConfigEM1D
    N, eps, dt, w, x, v = setup()
    rng = np.random.default_rng(1)
    y = rng.uniform(-0.01, 1.01, 5000)
    y[:3] = [0.001, 1.0 + 0.7 / N, 0.4 / N]          # wrapped-centre cases: no weight
    u = rng.standard_normal(5000)
    with pg.ContextEM1D(ConfigEM1D(N=N, P=1024, eps=eps, dt=dt, w=w)) as c:
        assert same(c.deposit(y, w), orc.em_deposit(N, y, w))
        assert same(c.deposit(y, -w, u), orc.em_deposit(N, y, -w, u))
        assert same(c.deposit(y[:3], w), np.zeros(N))


def test_fields_and_push_bitwise(pg, orc):
    from picgolf.core_em1d import ConfigEM1D
    N, eps, dt, w, x, v = setup()
    rng = np.random.default_rng(2)
    rho, Jy, Jz, Fx, Fy, Fz, Cy, Cz = (rng.standard_normal(N) for _ in range(8))
    E, B = rng.standard_normal((3, N)), rng.standard_normal((2, N))
    with pg.ContextEM1D(ConfigEM1D(N=N, P=1024, eps=eps, dt=dt, w=w)) as c:
        E1, B1, e3, b2 = c.fields(rho, Jy, Jz, Fx, Fy, Fz, Cy, Cz, E, B)
        o = orc.em_fields(eps, dt, rho, Jy, Jz, Fy, Fz, B[0], B[1], Cy, Cz)
        assert same(E1[0], o[0]) and same(E1[1], o[1]) and same(E1[2], o[2])
        assert same(B1[0], o[3]) and same(B1[1], o[4])
        assert same(e3[1], o[5]) and same(e3[2], o[6]) and same(b2[0], o[7]) and same(b2[1], o[8])
        assert same(e3[0], (o[0] + Fx) / 2)
        cc = rng.uniform(-0.01, 1.01, 2000)
        u = rng.standard_normal((3, 2000))
        vv = c.push(cc, e3, b2, u)
        ref = np.array([orc.em_push(cc[j], *e3, *b2, dt, u[:, j]) for j in range(2000)]).T
        assert same(vv, ref)


@pytest.mark.parametrize("P", [1024, 8192])
@pytest.mark.parametrize("sort", [1, 0])
def test_run_bitwise_over_two_calls(pg, orc, sort, P):
    """The scripts' P = 1024, and 8192 (eight deposit blocks per sweep)."""
    from picgolf.core_em1d import ConfigEM1D
    N, eps, dt, w, x, v = setup(P=P)
    T1, T2 = 5, 7
    with pg.ContextEM1D(ConfigEM1D(N=N, P=P, eps=eps, dt=dt, w=w, sort=bool(sort))) as c:
        c.upload(x, v)
        f1, s1 = c.run(T1)
        f2, s2 = c.run(T2)
        xg, vg, Eg, Bg = c.download()
    xo, vo, Eo, Bo, _, fo, so = orc.em_run(N, eps, dt, w, x, v, T1 + T2, sort=sort)
    assert np.array_equal(np.concatenate([s1, s2]), so) and np.all(so >= 2)
    assert same(np.concatenate([f1, f2], axis=1), fo)
    assert same(xg, xo) and same(vg, vo) and same(Eg, Eo) and same(Bg, Bo)


def test_run_without_fields_and_state_roundtrip(pg, orc):
    from picgolf.core_em1d import ConfigEM1D
    N, eps, dt, w, x, v = setup(seed=3)
    rng = np.random.default_rng(3)
    E, B = 0.1 * rng.standard_normal((3, N)), np.vstack([0.1 * rng.standard_normal(N), np.ones(N)])
    with pg.ContextEM1D(ConfigEM1D(N=N, P=1024, eps=eps, dt=dt, w=w)) as c:
        c.upload(x, v, E, B)
        xg, vg, Eg, Bg = c.download()
        assert same(xg, x) and same(vg, v) and same(Eg, E) and same(Bg, B)
        f, s = c.run(4, want_fields=False)
        assert f is None
        xg, vg, Eg, Bg = c.download()
    xo, vo, Eo, Bo, _, _, so = orc.em_run(N, eps, dt, w, x, v, 4, E=E, B=B, want_fields=False)
    assert np.array_equal(s, so)
    assert same(xg, xo) and same(vg, vo) and same(Eg, Eo) and same(Bg, Bo)


def test_sort_changes_no_field_bit_at_scale(pg, orc):
    """N = 1024, P = 2^20: the run with the per-step sort and the run without
    it give the same twelve fields bit for bit (every sum is exact), and the
    sorted particles are the unsorted ones in stable x order."""
    from picgolf.core_em1d import ConfigEM1D
    N, P = 1024, 1 << 20
    _, eps, dt, w, x, v = setup(N=N, P=P, seed=4)
    out = {}
    for sort in (True, False):
        with pg.ContextEM1D(ConfigEM1D(N=N, P=P, eps=eps, dt=dt, w=w, sort=sort)) as c:
            c.upload(x, v)
            f, s = c.run(3)
            out[sort] = (f, s, *c.download())
    assert same(out[True][0], out[False][0]) and np.array_equal(out[True][1], out[False][1])
    xs, vs = out[True][2], out[True][3]
    xu, vu = out[False][2], out[False][3]
    order = np.argsort(xu, kind="stable")
    assert np.all(np.diff(xs) >= 0)
    assert same(xs, xu[order]) and same(vs, vu[:, order])
