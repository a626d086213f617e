"""GPU parity of the Lorenz-gauge electromagnetic 2D3V path
(LorenzGaugeStaggeredField of src/PIC2D3V.jl as src/2D3V.jl drives it)
through the C ABI in include/pic2d3v.h: warmup!, steps of loop! with and
without diagnose! (rows, the 18 state arrays, the particles and the fourteen
binned fields) bit for bit against the CPU restatement, for every shape and
for explicit, semi-implicit and implicit theta; a step from an arbitrary
complex state; 20 steps across two regroupings of the held particles; a
128 x 128 run with 2 x 2^18 particles; LorenzGaugeField's
centred loop! (every shape tested, theta 0 and 1, from a fresh and from an
arbitrary state); and, at 512 x 512
with 2 x 2^22 particles, the size-independent properties that the fields do
not depend on the particle order (the LDS-window deposit against the global
fallback) and that two calls equal one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "c":
        a, b = a.view(np.float64), np.asarray(b, dtype=np.complex128).view(np.float64)
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def make_species(P, Lx, Ly, shape="bspline2", seed=0, sort_dxdy=None):
    rng = np.random.default_rng(seed)
    out = []
    for q, m, vth in ((-1.0, 1.0, 0.3), (1.0, 32.0, 0.05)):
        xyv = np.vstack([rng.uniform(0, Lx, P), rng.uniform(0, Ly, P), vth * rng.standard_normal((3, P))])
        out.append(dict(shape=shape, charge=q, mass=m, weight=Lx * Ly / P, xyv=xyv))
    return out


def context(pg, NX, NY, Lx, Ly, dt, B0, species, theta=1.0, buffer=3, ntskip=1, ngskip=1, field="lorenz_staggered"):
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    sc = [SpeciesConfig(P=s["xyv"].shape[1], shape=s["shape"], charge=s["charge"], mass=s["mass"],
                        weight=s["weight"]) for s in species]
    cfg = Config2D3V(NX=NX, NY=NY, dt=dt, species=sc, Lx=Lx, Ly=Ly, B0=tuple(B0), buffer=buffer, ntskip=ntskip,
                     ngskip=ngskip, field=field, theta=theta)
    c = pg.Context2D3V(cfg)
    for k, s in enumerate(species):
        c.upload(k, s["xyv"])
    return c


def check_state(c, species, F):
    for k, s in enumerate(species):
        assert same(c.download(k), s["xyv"]), f"species {k}"
    G = c.lg_state()
    for k in range(18):
        assert same(G[k], F[k]), f"state slot {k}"


@pytest.mark.parametrize("shape", ["bspline2", "ngp", "area", "bspline1", "bspline3", "bspline4", "bspline5"])
def test_warmup_and_steps_bitwise(pg, orc, shape):
    NX, NY, Lx, Ly, dt = 32, 16, 1.0, 2.0, 0.01
    B0 = (0.6, 0.0, 0.2)
    sp = make_species(2048, Lx, Ly, shape=shape, seed=1)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, buffer=3, ngskip=2) as c:
        c.warmup()
        sp1, F1 = orc.lg_warmup(NX, NY, Lx, Ly, dt, sp)
        check_state(c, sp1, F1)
        rows = c.run(3)
        sp2, F2, rows2, bins2 = orc.lg_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, 3, buffer=3, ngskip=2)
        check_state(c, sp2, F2)
        assert same(rows, rows2)
        assert same(c.lg_bins(), bins2)
        Exs, Eys, phis = c.fields()
        assert same(Exs, bins2[0]) and same(Eys, bins2[1]) and same(phis, bins2[9])


@pytest.mark.parametrize("theta,dt", [(0.0, 0.002), (0.5, 0.01), (1.0, 0.05)])
def test_loop_without_diagnostics_and_theta(pg, orc, theta, dt):
    NX, NY, Lx, Ly = 16, 32, 2.0, 1.0
    B0 = (0.0, 0.5, 0.0)
    sp = make_species(1024, Lx, Ly, seed=2)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=theta) as c:
        c.warmup()
        c.loop(4)
        sp1, F1 = orc.lg_warmup(NX, NY, Lx, Ly, dt, sp)
        sp2, F2, rows2, _ = orc.lg_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, 4, theta=theta, diagnose=False)
        check_state(c, sp2, F2)
        # then with diagnostics, from step t = 4, ntskip 1
        rows = c.run(2)
        sp3, F3, rows3, bins3 = orc.lg_run(NX, NY, Lx, Ly, dt, np.array(B0), sp2, F2, 2, theta=theta, t0=4,
                                           buffer=3)
        check_state(c, sp3, F3)
        assert same(rows, rows3)
        assert same(c.lg_bins(nbins=6)[:, 4:], bins3[:, 4:])
        assert not np.any(c.lg_bins(nbins=6)[:, :4])


def test_step_from_an_arbitrary_state(pg, orc):
    NX, NY, Lx, Ly, dt = 32, 32, 1.0, 1.0, 0.02
    B0 = (0.3, -0.2, 0.7)
    sp = make_species(512, Lx, Ly, seed=3)
    rng = np.random.default_rng(3)
    F = 0.1 * (rng.standard_normal((18, NY, NX)) + 1j * rng.standard_normal((18, NY, NX)))
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, buffer=2, ntskip=2, ngskip=4) as c:
        for k in range(18):
            c.lg_set(k, F[k])
        assert same(c.lg_state(), F)
        rows = c.run(3)
        sp1, F1, rows1, bins1 = orc.lg_run(NX, NY, Lx, Ly, dt, np.array(B0), sp, F, 3, buffer=2, ntskip=2, ngskip=4)
        check_state(c, sp1, F1)
        assert same(rows, rows1) and rows.shape == (2, 11)
        assert same(c.lg_bins(), bins1)


def test_regroupings_over_many_steps_bitwise(pg, orc):
    """The context regroups its particles by tile before the first step and
    every 16 steps; 20 diagnosed steps (two regroupings) match the oracle,
    which keeps the caller's order, in the rows, state, particles and bins."""
    NX, NY, Lx, Ly, dt = 32, 32, 1.0, 1.0, 0.02
    B0 = (0.5, 0.0, 0.3)
    sp = make_species(2048, Lx, Ly, seed=9)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, buffer=3) as c:
        c.warmup()
        rows = c.run(20)
        sp1, F1 = orc.lg_warmup(NX, NY, Lx, Ly, dt, sp)
        sp2, F2, rows2, bins2 = orc.lg_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, 20, buffer=3)
        check_state(c, sp2, F2)
        assert same(rows, rows2) and same(c.lg_bins(), bins2)


def test_mid_size_bitwise(pg, orc):
    NX, NY, Lx, Ly, dt = 128, 128, 1.0, 1.0, 1.0 / 1024
    B0 = (0.5, 0.0, 0.0)
    sp = make_species(1 << 18, Lx, Ly, seed=4)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, buffer=10) as c:
        for k in range(2):
            c.sort(k, Lx / NX, Ly / NY)
        sp = [dict(s, xyv=c.download(k)) for k, s in enumerate(sp)]
        c.warmup()
        rows = c.run(2)
        sp1, F1 = orc.lg_warmup(NX, NY, Lx, Ly, dt, sp)
        sp2, F2, rows2, bins2 = orc.lg_run(NX, NY, Lx, Ly, dt, np.array(B0), sp1, F1, 2, buffer=10)
        check_state(c, sp2, F2)
        assert same(rows, rows2)


def test_order_independence_and_split_calls_at_scale(pg):
    """512 x 512, 2 x 2^22 particles: sorted particles (LDS windows) and the
    same particles in random order (global fallback) give the same fields
    and particles bit for bit (the particle sums of the rows follow the
    order); two calls equal one."""
    NX, NY, Lx, Ly, dt = 512, 512, 1.0, 1.0, 1.0 / 4096
    B0 = (0.5, 0.0, 0.0)
    sp = make_species(1 << 22, Lx, Ly, seed=5)
    state = {}
    for order in ("sorted", "shuffled"):
        with context(pg, NX, NY, Lx, Ly, dt, B0, sp, buffer=10) as c:
            perms = []
            for k in range(2):
                perms.append(c.sort(k, Lx / NX, Ly / NY))
            if order == "shuffled":
                rng = np.random.default_rng(6)
                for k in range(2):
                    xyv = c.download(k)
                    q = rng.permutation(xyv.shape[1])
                    c.upload(k, xyv[:, q])
                    perms[k] = perms[k][q]
            c.warmup()
            rows = c.run(2)
            F = c.lg_state()
            xs = [c.download(k) for k in range(2)]
            state[order] = (rows, F, xs, perms)
    a, b = state["sorted"], state["shuffled"]
    # fields and the field-derived row entries are exact sums of the particles
    assert same(a[1], b[1])
    assert same(a[0][:, [1, 5, 6, 7]], b[0][:, [1, 5, 6, 7]])
    # the particle sums of the row follow the particles' order (canonical tree)
    assert np.allclose(a[0], b[0], rtol=1e-12, atol=1e-15)
    for k in range(2):
        # the same particles: undo both orders back to the upload order
        ia, ib = np.argsort(a[3][k]), np.argsort(b[3][k])
        assert same(a[2][k][:, ia], b[2][k][:, ib])
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, buffer=10) as c:
        for k in range(2):
            c.sort(k, Lx / NX, Ly / NY)
        c.warmup()
        r1 = c.run(1)
        r2 = c.run(1)
        assert same(np.concatenate([r1, r2]), a[0]) and same(c.lg_state(), a[1])


@pytest.mark.parametrize("shape,theta,dt", [("bspline2", 1.0, 0.02), ("area", 0.0, 0.002), ("bspline3", 0.5, 0.01),
                                            ("ngp", 1.0, 0.01)])
def test_centred_loop_bitwise(pg, orc, shape, theta, dt):
    """LorenzGaugeField: a fresh field (EBxyz zero until the first update!),
    then loop! over two calls."""
    NX, NY, Lx, Ly = 32, 16, 1.0, 2.0
    B0 = (0.4, 0.0, 0.3)
    sp = make_species(1536, Lx, Ly, shape=shape, seed=7)
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, theta=theta, field="lorenz") as c:
        c.warmup()                                   # warmup!(::AbstractField): nothing
        c.loop(2)
        c.loop(2)
        F0, EB0 = np.zeros((18, NY, NX), dtype=complex), np.zeros((6, NY, NX))
        sp1, F1, EB1 = orc.lg_run_centred(NX, NY, Lx, Ly, dt, np.array(B0), sp, F0, EB0, 4, theta=theta)
        check_state(c, sp1, F1)
        Ex, Ey = c.field()
        assert same(Ex, EB1[0]) and same(Ey, EB1[1])
        with pytest.raises(pg.PicError):
            c.run(1)


def test_centred_from_an_arbitrary_state(pg, orc):
    NX, NY, Lx, Ly, dt = 16, 32, 2.0, 1.0, 0.01
    B0 = (0.3, -0.2, 0.7)
    sp = make_species(512, Lx, Ly, seed=8)
    rng = np.random.default_rng(8)
    F = 0.1 * (rng.standard_normal((18, NY, NX)) + 1j * rng.standard_normal((18, NY, NX)))
    with context(pg, NX, NY, Lx, Ly, dt, B0, sp, field="lorenz") as c:
        for k in range(18):
            c.lg_set(k, F[k])
        c.loop(3)
        EB = orc.lg_update(F, np.array(B0))
        sp1, F1, EB1 = orc.lg_run_centred(NX, NY, Lx, Ly, dt, np.array(B0), sp, F, EB, 3, theta=1.0)
        check_state(c, sp1, F1)
        Ex, Ey = c.field()
        assert same(Ex, EB1[0]) and same(Ey, EB1[1])
