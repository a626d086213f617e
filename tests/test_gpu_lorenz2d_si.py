"""GPU parity of LorenzGaugeSemiImplicitField (src/PIC2D3V.jl:403-461,
warmup! :864-871, loop! :993-1100) through the C ABI in include/pic2d3v.h
(field PIC2D3V_LORENZ_SEMI_IMPLICIT): warmup! (the 22 state slots, the
buffered deposit grid and the particles) and steps of loop! with and without
diagnose! (rows, bins, iteration counts) bit for bit against the CPU
restatement, for every shape and for several field and source thetas; a step
from an arbitrary state and grid; the stopping rule (rtol = 0 runs
maxiters + 1 iterations); the BoundsError of a stencil outside the buffered
grid; a 64 x 64 run with 2 x 2^16 particles; and, at 256 x 256 with
2 x 2^20 particles, the size-independent properties that the state, the
grid and the iteration counts do not depend on the particle order and that
two calls equal one."""
import numpy as np
import pytest

from test_gpu_lorenz2d import make_species, same

pytestmark = pytest.mark.gpu

RTOL = float(np.sqrt(np.finfo(float).eps))


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def context(pg, NX, NY, Lx, Ly, dt, B0, species, theta=1.0, theta_source=0.0, rtol=RTOL, maxiters=10, buffer=3,
            ntskip=1, ngskip=1):
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    sc = [SpeciesConfig(P=s["xyv"].shape[1], shape=s["shape"], charge=s["charge"], mass=s["mass"],
                        weight=s["weight"]) for s in species]
    cfg = Config2D3V(NX=NX, NY=NY, dt=dt, species=sc, Lx=Lx, Ly=Ly, B0=tuple(B0), buffer=buffer, ntskip=ntskip,
                     ngskip=ngskip, field="lorenz_semi_implicit", theta=theta, theta_source=theta_source,
                     rtol=rtol, maxiters=maxiters)
    c = pg.Context2D3V(cfg)
    for k, s in enumerate(species):
        c.upload(k, s["xyv"])
    return c


def check(c, species, F, grid, EB=None):
    for k, s in enumerate(species):
        assert same(c.download(k), s["xyv"]), f"species {k}"
    G = c.lg_state()
    for k in range(22):
        assert same(G[k], F[k]), f"state slot {k}"
    assert same(c.lg_grid(), grid)
    if EB is not None:
        Ex, Ey = c.field()
        assert same(Ex, EB[0]) and same(Ey, EB[1])


@pytest.mark.parametrize("shape", ["bspline2", "ngp", "area", "bspline1", "bspline3", "bspline4", "bspline5"])
def test_warmup_and_steps_bitwise(pg, orc, shape):
    NX, NY, Lx, Ly, dt, b = 32, 16, 1.0, 2.0, 0.01, 3
    B0 = (0.6, 0.0, 0.2)
    sp = make_species(2048, Lx, Ly, shape=shape, seed=1)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=1.0, theta_source=0.05, buffer=b, ngskip=2) as c:
        c.warmup()
        sp1, F1, g1 = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp, b)
        check(c, sp1, F1, g1)
        rows = c.run(3)
        o = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, g1, np.zeros((6, NY, NX)), 3, theta=1.0,
                          theta_source=0.05, buffer=b, ngskip=2)
        check(c, o[0], o[1], o[2], o[3])
        assert same(rows, o[4]) and same(c.lg_bins(), o[5])
        assert np.array_equal(c.lg_iterations(), o[6]) and np.all(o[6] > 1)


@pytest.mark.parametrize("theta,theta_source,dt", [(0.0, 0.0, 0.002), (0.5, 0.3, 0.01), (1.0, 1.0, 0.05)])
def test_loop_without_diagnostics_and_thetas(pg, orc, theta, theta_source, dt):
    NX, NY, Lx, Ly, b = 16, 32, 2.0, 1.0, 3
    B0 = (0.0, 0.5, 0.0)
    sp = make_species(1024, Lx, Ly, seed=2)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=theta, theta_source=theta_source, buffer=b) as c:
        c.warmup()
        c.loop(4)
        it = c.lg_iterations()
        sp1, F1, g1 = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp, b)
        o = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, g1, np.zeros((6, NY, NX)), 4, theta=theta,
                          theta_source=theta_source, buffer=b, diagnose=False)
        check(c, o[0], o[1], o[2], o[3])
        assert np.array_equal(it, o[6])
        # then with diagnostics from t = 4
        rows = c.run(2)
        p = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), o[0], o[1], o[2], o[3], 2, theta=theta,
                          theta_source=theta_source, buffer=b, t0=4)
        check(c, p[0], p[1], p[2], p[3])
        assert same(rows, p[4]) and np.array_equal(c.lg_iterations(), p[6])
        assert same(c.lg_bins(nbins=6)[:, 4:], p[5][:, 4:])


def test_regroupings_over_many_steps_bitwise(pg, orc):
    """The context regroups its particles (and rotates the buffers that hold
    them) before the first step and every 16 steps: 20 diagnosed steps of
    run(), and 17 steps of loop() followed by 3 of run(), cross a mid-run
    regroup and match the oracle, which keeps the caller's order, in the
    rows, state, grid, particles, bins and iteration counts."""
    NX, NY, Lx, Ly, dt, b = 32, 32, 1.0, 1.0, 0.02, 3
    B0 = (0.5, 0.0, 0.3)
    sp = make_species(2048, Lx, Ly, seed=9)
    sp1, F1, g1 = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp, b)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=0.8, theta_source=0.1, buffer=b) as c:
        c.warmup()
        rows = c.run(20)
        o = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, g1, np.zeros((6, NY, NX)), 20, theta=0.8,
                          theta_source=0.1, buffer=b)
        check(c, o[0], o[1], o[2], o[3])
        assert same(rows, o[4]) and same(c.lg_bins(), o[5])
        assert np.array_equal(c.lg_iterations(), o[6])
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=0.8, theta_source=0.1, buffer=b) as c:
        c.warmup()
        c.loop(17)
        o = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, g1, np.zeros((6, NY, NX)), 17, theta=0.8,
                          theta_source=0.1, buffer=b, diagnose=False)
        check(c, o[0], o[1], o[2], o[3])
        assert np.array_equal(c.lg_iterations(), o[6])
        rows = c.run(3)
        p = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), o[0], o[1], o[2], o[3], 3, theta=0.8,
                          theta_source=0.1, buffer=b, t0=17)
        check(c, p[0], p[1], p[2], p[3])
        assert same(rows, p[4]) and np.array_equal(c.lg_iterations(), p[6])


def test_step_from_an_arbitrary_state(pg, orc):
    NX, NY, Lx, Ly, dt, b = 32, 32, 1.0, 1.0, 0.02, 4
    B0 = (0.3, -0.2, 0.7)
    sp = make_species(512, Lx, Ly, seed=3)
    rng = np.random.default_rng(3)
    F = 0.1 * (rng.standard_normal((22, NY, NX)) + 1j * rng.standard_normal((22, NY, NX)))
    grid = 0.1 * rng.standard_normal((NY + 2 * b, NX + 2 * b, 4))
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=0.7, theta_source=0.2, buffer=b, ntskip=2, ngskip=4) as c:
        for k in range(22):
            c.lg_set(k, F[k])
        c.lg_set_grid(grid)
        assert same(c.lg_state(), F) and same(c.lg_grid(), grid)
        EB = orc.lg_update(F, np.array(B0))
        rows = c.run(3)
        o = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), sp, F, grid, EB, 3, theta=0.7, theta_source=0.2,
                          buffer=b, ntskip=2, ngskip=4)
        check(c, o[0], o[1], o[2], o[3])
        assert same(rows, o[4]) and rows.shape == (2, 11)
        assert same(c.lg_bins(), o[5]) and np.array_equal(c.lg_iterations(), o[6])


def test_stopping_rule(pg, orc):
    NX, NY, Lx, Ly, dt, b = 16, 16, 1.0, 1.0, 0.02, 3
    B0 = (0.0, 0.0, 1.0)
    sp = make_species(768, Lx, Ly, seed=4)
    sp1, F1, g1 = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp, b)
    for maxiters in (0, 2):
        with context(pg, NX, NY, Lx, Ly, dt, B0, sp, rtol=0.0, maxiters=maxiters, buffer=b) as c:
            c.warmup()
            c.loop(2)
            assert np.array_equal(c.lg_iterations(), [maxiters + 1] * 2)
            o = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, g1, np.zeros((6, NY, NX)), 2, rtol=0.0,
                              maxiters=maxiters, buffer=b, diagnose=False)
            check(c, o[0], o[1], o[2], o[3])


def test_stencil_outside_the_buffer_is_an_error(pg, orc):
    """A particle more than a box length from the box after warmup!'s move
    reaches past the buffered grid: the reference's BoundsError."""
    NX, NY, Lx, Ly, dt, b = 16, 16, 1.0, 1.0, 0.01, 3
    sp = make_species(256, Lx, Ly, seed=5)
    sp[0]["xyv"][2, 7] = 500.0 * Lx / dt
    with pytest.raises(ValueError):
        orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp, b)
    with context(pg, NX, NY, Lx, Ly, dt, (0.0, 0.0, 0.0), sp, buffer=b) as c:
        with pytest.raises(pg.PicError) as e:
            c.warmup()
        assert "BoundsError" in str(e.value)


def test_mid_size_bitwise(pg, orc):
    NX, NY, Lx, Ly, dt, b = 64, 64, 1.0, 1.0, 1.0 / 512, 10
    B0 = (0.5, 0.0, 0.0)
    sp = make_species(1 << 16, Lx, Ly, seed=6)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=1.0, theta_source=0.05, buffer=b) as c:
        for k in range(2):
            c.sort(k, Lx / NX, Ly / NY)
        sp = [dict(s, xyv=c.download(k)) for k, s in enumerate(sp)]
        c.warmup()
        rows = c.run(2)
        sp1, F1, g1 = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp, b)
        o = orc.lg_si_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, g1, np.zeros((6, NY, NX)), 2, theta=1.0,
                          theta_source=0.05, buffer=b)
        check(c, o[0], o[1], o[2], o[3])
        assert same(rows, o[4]) and np.array_equal(c.lg_iterations(), o[6])


def test_order_independence_and_split_calls_at_scale(pg):
    """256 x 256, 2 x 2^20 particles: sorted particles (LDS windows) and the
    same particles in random order (global fallback) give the same state,
    grid, iteration counts and particles bit for bit; two calls equal one."""
    NX, NY, Lx, Ly, dt, b = 256, 256, 1.0, 1.0, 1.0 / 2048, 10
    B0 = (0.5, 0.0, 0.0)
    sp = make_species(1 << 20, Lx, Ly, seed=7)
    state = {}
    for order in ("sorted", "shuffled"):
        with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta_source=0.05, buffer=b) as c:
            perms = [c.sort(k, Lx / NX, Ly / NY) for k in range(2)]
            if order == "shuffled":
                rng = np.random.default_rng(8)
                for k in range(2):
                    xyv = c.download(k)
                    q = rng.permutation(xyv.shape[1])
                    c.upload(k, xyv[:, q])
                    perms[k] = perms[k][q]
            c.warmup()
            rows = c.run(2)
            state[order] = (rows, c.lg_state(), c.lg_grid(), [c.download(k) for k in range(2)], perms,
                            c.lg_iterations())
    a, z = state["sorted"], state["shuffled"]
    assert same(a[1], z[1]) and same(a[2], z[2]) and np.array_equal(a[5], z[5])
    assert same(a[0][:, [1, 5, 6, 7]], z[0][:, [1, 5, 6, 7]])
    assert np.allclose(a[0], z[0], rtol=1e-12, atol=1e-15)
    for k in range(2):
        ia, iz = np.argsort(a[4][k]), np.argsort(z[4][k])
        assert same(a[3][k][:, ia], z[3][k][:, iz])
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta_source=0.05, buffer=b) as c:
        for k in range(2):
            c.sort(k, Lx / NX, Ly / NY)
        c.warmup()
        r1 = c.run(1)
        i1 = c.lg_iterations()
        r2 = c.run(1)
        assert same(np.concatenate([r1, r2]), a[0]) and same(c.lg_state(), a[1]) and same(c.lg_grid(), a[2])
        assert np.array_equal(np.concatenate([i1, c.lg_iterations()]), a[5])
