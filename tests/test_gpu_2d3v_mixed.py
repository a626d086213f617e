"""The mixed-precision electrostatic path of the structured 2D3V module
(loop! of src/PIC2D3V.jl:537-555 with precision = PIC_PREC_MIXED in
include/pic2d3v.h) through the C ABI, against its FP64 path (which is bit for
bit the CPU restatement, tests/test_gpu_2d3v.py) and against exact
properties of its own:

- the upload/download round trip: positions to half a 2^-32 quantum of the
  box (modulo the box: a position that rounds to zero is the box's edge),
  velocities to the rounding of their FP32 hi + lo pair;
- the charge is an exact integer sum and each particle's update its own, so
  the particles uploaded in another order, or sorted in the middle of a run,
  give bit-identical fields and (permuted back) particles;
- sort! permutes the caller's order as the restatement's sort of the
  downloaded state does, and moves the mixed state exactly;
- two-species runs over all eight shapes, both Boris variants and both field
  updates: diagnostics rows and the bulk of the particles within fixed bars
  of the FP64 run, and the largest field and particle differences within
  twice those of an FP64 run from positions displaced by a few quanta of the
  mixed resolution (the reference's shapes move a particle's charge by two
  cells when it crosses a cell edge, so any position difference at an edge
  shows as a jump);
- the momentum rows with the field replaced each step stay constant to the
  rounding of the FP32 gather (deposit and gather share the integer weights);
- 256 x 256 cells, 2^21 particles per species, 20 steps against FP64.
"""
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


def plasma(rng, P, Lx, Ly, shapes=("bspline2", "area"), vth=0.3):
    """Two species (electrons and heavier ions) as (SpeciesConfig, xyv) pairs."""
    from picgolf.core2d3v import SpeciesConfig
    out = []
    for k, (q, m) in enumerate(((-1.0, 1.0), (1.0, 16.0))):
        xyv = np.vstack([rng.uniform(0, Lx, P), rng.uniform(0, Ly, P), vth * rng.standard_normal((3, P))])
        out.append((SpeciesConfig(P=P, shape=shapes[k], charge=q, mass=m, weight=Lx * Ly / P), xyv))
    return out


def config(pl, **kw):
    from picgolf.core2d3v import Config2D3V
    base = dict(NX=32, NY=16, Lx=2.0, Ly=0.75, dt=0.02, B0=(0.2, -0.1, 1.5), buffer=3, ntskip=1, ngskip=2,
                precision="mixed")
    base.update(kw)
    return Config2D3V(species=[c for c, _ in pl], **base)


def run(pg, cfg, states, T):
    with pg.Context2D3V(cfg) as c:
        for s, xyv in enumerate(states):
            c.upload(s, xyv)
        rows = c.run(T)
        bins = c.fields()
        E = c.field()
        parts = [c.download(s) for s in range(len(states))]
    return rows, bins, E, parts


def periodic_diff(a, b, L):
    return np.abs((a - b + L / 2) % L - L / 2)


def test_upload_download_roundtrip(pg):
    rng = np.random.default_rng(21)
    pl = plasma(rng, 4000, 2.0, 0.75)
    xyv = pl[0][1]
    xyv[0, :4] = [2.0, 1.0, 1e-12, 2.0 - 1e-13]
    xyv[2, :4] = [0.0, -0.0, 1e-30, 3.0e5]
    with pg.Context2D3V(config(pl)) as c:
        c.upload(0, xyv)
        got = c.download(0)
    for k, L in ((0, 2.0), (1, 0.75)):
        assert np.all((got[k] > 0) & (got[k] <= L))
        assert periodic_diff(got[k], xyv[k], L).max() <= L * 2.0 ** -33 * (1 + 1e-12)
    assert got[0, 0] == 2.0 and got[0, 1] == 1.0
    v, w = got[2:], xyv[2:]
    assert np.all(np.abs(v - w) <= 2.0 ** -47 * np.abs(w))


def test_particle_order_and_sort_change_no_field_bit(pg):
    rng = np.random.default_rng(22)
    pl = plasma(rng, 6000, 2.0, 0.75, shapes=("bspline3", "area"))
    cfg = config(pl)
    states = [x for _, x in pl]
    rows, bins, E, parts = run(pg, cfg, states, 40)
    perms = [rng.permutation(x.shape[1]) for x in states]
    rows2, bins2, E2, parts2 = run(pg, cfg, [x[:, p] for x, p in zip(states, perms)], 40)
    assert same(E[0], E2[0]) and same(E[1], E2[1])
    for k in range(3):
        assert same(bins[k], bins2[k])
    assert same(rows[:, 1], rows2[:, 1])                   # field energy: from the field alone
    np.testing.assert_allclose(rows2, rows, rtol=1e-12, atol=1e-12 * np.abs(rows[:, 5:]).max())  # sums in another order
    for s in range(2):
        assert same(parts2[s], parts[s][:, perms[s]])
    # sort! in the middle of the run: the same fields bit for bit
    with pg.Context2D3V(cfg) as c:
        for s, xyv in enumerate(states):
            c.upload(s, xyv)
        c.run(17)
        perm = c.sort(0, 2.0 / 32, 0.75 / 16)
        c.run(23)
        E3 = c.field()
        p3 = c.download(0)
    assert same(E3[0], E[0]) and same(E3[1], E[1])
    assert same(p3, parts[0][:, perm])


def test_sort_follows_the_downloaded_state(pg, orc):
    rng = np.random.default_rng(23)
    pl = plasma(rng, 20000, 2.0, 0.75)
    with pg.Context2D3V(config(pl)) as c:
        for s, (_, xyv) in enumerate(pl):
            c.upload(s, xyv)
        c.run(3)
        before = c.download(0)
        perm = c.sort(0, 2.0 / 32, 0.75 / 16)
        got = c.download(0)
    ref, operm = orc.sort23(before, 2.0 / 32, 0.75 / 16)
    assert np.array_equal(perm, operm)
    assert same(got, ref) and same(got, before[:, perm])


def errors(a, b, cfg):
    """Relative errors of run a against run b: field energy, kinetic energy,
    momentum (over the characteristic momentum); the gathered field (largest
    and rms difference over the largest and rms field) and the binned fields;
    positions (in cells) and velocities (over the largest speed), largest
    and 99.9th percentile over the particles."""
    ra, ba, Ea, pa = a
    rb, bb, Eb, pb = b
    out = dict(
        field_energy=np.max(np.abs(ra[:, 1] - rb[:, 1]) / np.abs(rb[:, 1])),
        kinetic=np.max(np.abs(ra[:, 0] - rb[:, 0]) / np.abs(rb[:, 0])),
        momentum=np.max(np.abs(ra[:, 2:5] - rb[:, 2:5])) / np.abs(rb[:, 5:8]).max(),
        field=max(np.abs(Ea[k] - Eb[k]).max() / np.abs(Eb[k]).max() for k in range(2)),
        field_rms=max(np.sqrt(np.mean((Ea[k] - Eb[k]) ** 2) / np.mean(Eb[k] ** 2)) for k in range(2)),
        bins=max(np.abs(ba[k] - bb[k]).max() / np.abs(bb[k]).max() for k in range(3)),
    )
    dx = np.concatenate([np.maximum(periodic_diff(pa[s][0], pb[s][0], cfg.Lx) * cfg.NX / cfg.Lx,
                                    periodic_diff(pa[s][1], pb[s][1], cfg.Ly) * cfg.NY / cfg.Ly)
                         for s in range(len(pa))])
    dv = np.concatenate([np.abs(pa[s][2:] - pb[s][2:]).max(axis=0) / np.abs(pb[s][2:]).max() for s in range(len(pa))])
    out.update(position_cells=dx.max(), velocity=dv.max(), position_q999=np.quantile(dx, 0.999),
               velocity_q999=np.quantile(dv, 0.999))
    return out


def displaced(states, rng, cfg, quanta=4):
    """The states with every position moved by up to `quanta` 2^-32 fractions
    of the box (the mixed form's position resolution) and wrapped into (0, L]."""
    out = []
    for x in states:
        y = x.copy()
        for k, L in ((0, cfg.Lx), (1, cfg.Ly)):
            y[k] = y[k] + rng.uniform(-quanta, quanta, y.shape[1]) * L * 2.0 ** -32
            y[k] = np.where(y[k] > L, y[k] - L, np.where(y[k] > 0, y[k], y[k] + L))
        out.append(y)
    return out


# The reference's shapes (all but NGP) take r = ceil(zN) - zN from the lower
# cell edge, so a particle's charge moves by two cells when it crosses an
# edge: two runs whose positions differ by dz cells deposit a particle two
# cells apart with probability about 2 dz per step.  Field and particle
# errors that such events set are held to twice those of an FP64 run whose
# positions start displaced by a few quanta of the mixed resolution (or to
# the floor given); the global diagnostics and the bulk of the particles
# (99.9th percentile) to fixed bars.
EVENT = dict(field=1e-5, field_rms=5e-6, bins=1e-5, position_cells=1e-5, velocity=1e-5)
FIXED = dict(field_energy=1e-5, kinetic=1e-8, momentum=1e-8, position_q999=2e-5, velocity_q999=2e-6)


def check(em, ec, what):
    print(f"{what}: mixed vs fp64 " + " ".join(f"{k}={v:.2e}" for k, v in em.items()))
    print(f"{what}: displaced fp64 control " + " ".join(f"{k}={v:.2e}" for k, v in ec.items()))
    for k, bar in FIXED.items():
        assert em[k] <= bar, (k, em[k])
    for k, floor in EVENT.items():
        assert em[k] <= max(floor, 2 * ec[k]), (k, em[k], ec[k])


# (shapes of the two species, field update, Boris variant): every shape once
CASES = [
    (("ngp", "area"), "accumulate", "reference"),
    (("bspline0", "bspline1"), "replace", "textbook"),
    (("bspline2", "bspline3"), "accumulate", "textbook"),
    (("bspline4", "bspline5"), "replace", "reference"),
]


@pytest.mark.parametrize("shapes,field_update,boris", CASES)
def test_run_against_fp64(pg, shapes, field_update, boris):
    rng = np.random.default_rng(24)
    pl = plasma(rng, 4000, 2.0, 0.75, shapes=shapes)
    states = [x for _, x in pl]
    T = 100
    cm = config(pl, field_update=field_update, boris=boris)
    cf = config(pl, field_update=field_update, boris=boris, precision="fp64")
    fp64 = run(pg, cf, states, T)
    em = errors(run(pg, cm, states, T), fp64, cf)
    ec = errors(run(pg, cf, displaced(states, rng, cf), T), fp64, cf)
    check(em, ec, f"{shapes} {field_update} {boris}")


def test_momentum_with_the_field_replaced(pg):
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    rng = np.random.default_rng(25)
    P = 1 << 18
    xyv = np.vstack([rng.uniform(0, 1, P), rng.uniform(0, 1, P), 0.05 * rng.standard_normal((3, P))])
    for shape in ("area", "bspline2", "bspline5"):
        sc = SpeciesConfig(P=P, shape=shape, charge=-1.0, mass=1.0, weight=1.0 / P)
        cfg = Config2D3V(NX=128, NY=128, dt=0.01, species=[sc], field_update="replace", ntskip=1, ngskip=4,
                         precision="mixed")
        with pg.Context2D3V(cfg) as c:
            c.upload(0, xyv)
            rows = c.run(30)
        drift = np.abs(rows[:, 2:5] - rows[0, 2:5]).max() / rows[0, 5:8].max()
        print(f"{shape}: momentum drift {drift:.2e} of the characteristic momentum")
        assert drift <= 1e-6


def test_large_grid_against_fp64(pg):
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    rng = np.random.default_rng(26)
    P = 1 << 21
    states, sps = [], []
    for q, m, shape in ((-1.0, 1.0, "bspline2"), (1.0, 100.0, "bspline2")):
        states.append(np.vstack([rng.uniform(0, 1, P), rng.uniform(0, 1, P), 0.1 * rng.standard_normal((3, P))]))
        sps.append(SpeciesConfig(P=P, shape=shape, charge=q, mass=m, weight=1.0 / P))
    kw = dict(NX=256, NY=256, dt=0.01, species=sps, B0=(0.0, 0.0, 0.5), ntskip=1, ngskip=4)
    cf = Config2D3V(precision="fp64", **kw)
    fp64 = run(pg, cf, states, 20)
    em = errors(run(pg, Config2D3V(precision="mixed", **kw), states, 20), fp64, cf)
    ec = errors(run(pg, cf, displaced(states, rng, cf), 20), fp64, cf)
    check(em, ec, "256x256, 2 x 2^21")
