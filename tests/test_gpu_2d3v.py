"""GPU parity of the electrostatic path of the structured 2D3V module
(src/PIC2D3V.jl) through the C ABI in include/pic2d3v.h: the operators for
every shape and both Boris variants, whole runs of two species (rows, binned
fields, particles and the gathered field bit for bit against the CPU
restatement, split over calls), the per-cell sort's permutation, the
non-finite report, and size-independent properties at a 256 x 256 grid with
2^21 particles (the sort leaves the field bit for bit unchanged; momentum is
conserved with the field replaced each step)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = ["ngp", "area"] + [f"bspline{n}" for n in range(6)]


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def plasma(rng, P, Lx, Ly, shapes=("bspline2", "area"), vth=0.3):
    """Two species (electrons and heavier ions) as (SpeciesConfig, xyv) pairs."""
    from picgolf.core2d3v import SpeciesConfig
    out = []
    for k, (q, m) in enumerate(((-1.0, 1.0), (1.0, 16.0))):
        xyv = np.vstack([rng.uniform(0, Lx, P), rng.uniform(0, Ly, P), vth * rng.standard_normal((3, P))])
        out.append((SpeciesConfig(P=P, shape=shapes[k], charge=q, mass=m, weight=Lx * Ly / P), xyv))
    return out


def oracle_species(pl):
    return [dict(shape=c.shape, charge=c.charge, mass=c.mass, weight=c.weight, xyv=x) for c, x in pl]


def config(pl, **kw):
    from picgolf.core2d3v import Config2D3V
    base = dict(NX=32, NY=16, Lx=2.0, Ly=0.75, dt=0.02, B0=(0.2, -0.1, 1.5), buffer=3, ntskip=3, ngskip=2)
    base.update(kw)
    return Config2D3V(species=[c for c, _ in pl], **base)


@pytest.mark.parametrize("shape", SHAPES)
def test_deposit_and_gather_bitwise(pg, orc, shape):
    rng = np.random.default_rng(1)
    pl = plasma(rng, 64, 2.0, 0.75)
    cfg = config(pl)
    n = 5000
    x, y = rng.uniform(0, 2.0, n), rng.uniform(0, 0.75, n)
    x[:4] = [2.0, 1.0, 0.0625, 1e-12]                        # top edge, exact cell edges, tiny
    with pg.Context2D3V(cfg) as c:
        for w in (-0.37, 2.5):
            assert same(c.deposit(shape, x, y, w), orc.deposit23(32, 16, x, y, w, shape, 2.0, 0.75))
        Ex, Ey = rng.standard_normal((16, 32)), rng.standard_normal((16, 32))
        ex, ey = c.gather(shape, Ex, Ey, x, y)
        ref = np.array([orc.gather23(Ex, Ey, x[j], y[j], shape, 2.0, 0.75) for j in range(n)])
        assert same(ex, ref[:, 0]) and same(ey, ref[:, 1])


@pytest.mark.parametrize("variant", ["reference", "textbook"])
def test_boris_bitwise(pg, orc, variant):
    rng = np.random.default_rng(2)
    pl = plasma(rng, 64, 2.0, 0.75)
    cfg = config(pl, boris=variant)
    n = 3000
    vx, vy, vz = rng.standard_normal((3, n))
    ex, ey = rng.standard_normal((2, n))
    with pg.Context2D3V(cfg) as c:
        for q_m in (-1.0, 1.0 / 16, 3.0):
            a, b, d = vx.copy(), vy.copy(), vz.copy()
            c.boris(a, b, d, ex, ey, q_m)
            ref = np.array([orc.boris23(vx[j], vy[j], vz[j], ex[j], ey[j], cfg.B0, cfg.dt, q_m,
                                        variant=0 if variant == "reference" else 1) for j in range(n)])
            assert same(a, ref[:, 0]) and same(b, ref[:, 1]) and same(d, ref[:, 2])


def test_fieldsolve_bitwise_and_nonfinite(pg, orc):
    rng = np.random.default_rng(3)
    pl = plasma(rng, 64, 2.0, 0.75)
    with pg.Context2D3V(config(pl)) as c:
        rho = rng.standard_normal((16, 32))
        Ex, Ey, phi = c.fieldsolve(rho)
        Eo, Fo, Po, bad = orc.fieldsolve23(rho, 2.0, 0.75)
        assert not bad and same(Ex, Eo) and same(Ey, Fo) and same(phi, Po)
        rho[3, 5] = np.nan
        with pytest.raises(pg.PicError) as e:
            c.fieldsolve(rho)
        assert e.value.code == -4


@pytest.mark.parametrize("field_update,boris", [("accumulate", "reference"), ("replace", "textbook")])
def test_run_two_species_bitwise(pg, orc, field_update, boris):
    rng = np.random.default_rng(4)
    pl = plasma(rng, 3000, 2.0, 0.75, shapes=("bspline3", "ngp"))
    cfg = config(pl, field_update=field_update, boris=boris)
    T1, T2 = 7, 11
    with pg.Context2D3V(cfg) as c:
        for s, (_, xyv) in enumerate(pl):
            c.upload(s, xyv)
        rows = np.vstack([c.run(T1), c.run(T2)])
        bins = c.fields()
        Eacc = c.field()
        got = [c.download(s) for s in range(2)]
    sp, (Ex, Ey), orow, obins, bad = orc.run23(
        32, 16, 2.0, 0.75, cfg.dt, cfg.B0, oracle_species(pl), T1 + T2, buffer=3, ntskip=3, ngskip=2,
        accumulate=1 if field_update == "accumulate" else 0, boris=0 if boris == "reference" else 1, nthreads=8)
    assert not bad
    assert rows.shape == (6, 8) and same(rows, orow)
    for k in range(3):
        assert same(bins[k], obins[k])
    assert same(Eacc[0], Ex) and same(Eacc[1], Ey)
    for s in range(2):
        assert same(got[s], sp[s]["xyv"])


def test_run_every_shape_one_step_each(pg, orc):
    """Two steps of one species for every shape (the push kernels of all
    eight instantiations), diagnostics every step, no subsampling."""
    from picgolf.core2d3v import SpeciesConfig
    rng = np.random.default_rng(5)
    for shape in SHAPES:
        xyv = np.vstack([rng.uniform(0, 1, 2000), rng.uniform(0, 1, 2000), 0.5 * rng.standard_normal((3, 2000))])
        sc = SpeciesConfig(P=2000, shape=shape, charge=-1.0, mass=1.0, weight=1 / 2000)
        from picgolf.core2d3v import Config2D3V
        cfg = Config2D3V(NX=16, NY=16, dt=0.05, species=[sc], B0=(0.0, 0.0, 2.0), ntskip=1, ngskip=1)
        with pg.Context2D3V(cfg) as c:
            c.upload(0, xyv)
            rows = c.run(2)
            got = c.download(0)
        sp, _, orow, _, _ = orc.run23(16, 16, 1.0, 1.0, 0.05, cfg.B0, [dict(shape=shape, charge=-1.0, mass=1.0,
                                                                              weight=1 / 2000, xyv=xyv)], 2)
        assert same(rows, orow), shape
        assert same(got, sp[0]["xyv"]), shape


def test_sort_bitwise(pg, orc):
    rng = np.random.default_rng(6)
    pl = plasma(rng, 20000, 2.0, 0.75)
    xyv = pl[0][1]
    xyv[2] = rng.choice([-1.0, -0.0, 0.0, 0.25, 1.0, np.nan], xyv.shape[1])     # ties, signed zeros, NaN
    xyv[0, :5] = 2.0
    with pg.Context2D3V(config(pl)) as c:
        c.upload(0, xyv)
        perm = c.sort(0, 2.0 / 32, 0.75 / 16)
        got = c.download(0)
        # a sort of too coarse cells is refused before any bucket is ordered
        with pytest.raises(pg.PicError) as e:
            c.sort(0, 100.0, 100.0)
        assert e.value.code == -1 and "share one cell" in str(e.value)
    ref, operm = orc.sort23(xyv, 2.0 / 32, 0.75 / 16)
    assert np.array_equal(perm, operm)
    assert same(got, ref)


def test_sort_after_steps_follows_the_caller_order(pg, orc):
    """The context regroups its particles by tile while it steps; sort!
    afterwards permutes the caller's order (the downloaded one), and a second
    run continues from the sorted order: its rows (whose particle sums follow
    that order), gathered field and particles match the oracle's run from
    the same state."""
    rng = np.random.default_rng(16)
    pl = plasma(rng, 20000, 2.0, 0.75)
    cfg = config(pl)
    with pg.Context2D3V(cfg) as c:
        for s, (_, xyv) in enumerate(pl):
            c.upload(s, xyv)
        c.run(3)
        before = c.download(0)
        perm = c.sort(0, 2.0 / 32, 0.75 / 16)
        got = c.download(0)
        ref, operm = orc.sort23(before, 2.0 / 32, 0.75 / 16)
        assert np.array_equal(perm, operm) and same(got, ref)
        other = c.download(1)
        Eacc = c.field()
        r = c.run(4)
        after = [c.download(s) for s in range(2)]
        Eacc2 = c.field()
    sp = oracle_species(pl)
    sp[0]["xyv"], sp[1]["xyv"] = got, other
    sp2, (Ex, Ey), orow, _, bad = orc.run23(32, 16, 2.0, 0.75, cfg.dt, cfg.B0, sp, 4, t0=3, buffer=3, ntskip=3,
                                            ngskip=2, Eacc=Eacc, nthreads=8)
    assert not bad
    assert r.shape == orow.shape == (2, 8) and same(r, orow)
    assert same(Eacc2[0], Ex) and same(Eacc2[1], Ey)
    for s in range(2):
        assert same(after[s], sp2[s]["xyv"])


def test_upload_download_roundtrip_and_reset(pg, orc):
    rng = np.random.default_rng(7)
    pl = plasma(rng, 1000, 2.0, 0.75)
    with pg.Context2D3V(config(pl)) as c:
        for s, (_, xyv) in enumerate(pl):
            c.upload(s, xyv)
            assert same(c.download(s), xyv)
        r1 = c.run(5)
        E1 = c.field()
        for s, (_, xyv) in enumerate(pl):
            c.upload(s, xyv)
        c.reset()                                   # fresh field, t = 0, no bins
        assert not np.any(c.field()[0])
        r2 = c.run(5)
        assert same(r1, r2) and same(E1[0], c.field()[0])
    with pg.Context2D3V(config(pl)) as c:
        with pytest.raises(pg.PicError) as e:
            c.run(1)
        assert e.value.code == -5


def test_large_grid_properties(pg):
    """256 x 256 cells, 2^21 particles of one species (area shape), B = 0, the
    field replaced each step: (a) a per-cell sort changes no field bit (the
    charge is an exact sum, each particle's push its own); (b) the momentum
    rows stay constant to roundoff (the deposit and the gather are one
    shape's transpose, the spectral gradient is odd in k)."""
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    rng = np.random.default_rng(8)
    P = 1 << 21
    xyv = np.vstack([rng.uniform(0, 1, P), rng.uniform(0, 1, P), 0.05 * rng.standard_normal((3, P))])
    sc = SpeciesConfig(P=P, shape="area", charge=-1.0, mass=1.0, weight=1.0 / P)
    cfg = Config2D3V(NX=256, NY=256, dt=0.01, species=[sc], field_update="replace", ntskip=1, ngskip=4)
    with pg.Context2D3V(cfg) as c:
        c.upload(0, xyv)
        rows = c.run(20)
        Ea = c.field()
        c.reset()
        c.upload(0, xyv)
        perm = c.sort(0, 1 / 256, 1 / 256)
        assert np.array_equal(np.sort(perm), np.arange(P))
        c.run(20)
        Eb = c.field()
    assert same(Ea[0], Eb[0]) and same(Ea[1], Eb[1])
    drift = np.abs(rows[:, 2:5] - rows[0, 2:5]).max()
    assert drift <= 1e-12 * rows[0, 5:8].max()
