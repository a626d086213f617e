"""The mixed-precision form of the Simpson-1/3 fixed point
(src/GaussianFixedPointQuietSimpson13.jl, src/AreaFixedPointQuietSimpson13.jl)
against its FP64 path, at the bars the mixed midpoint is held to: field-energy
relative error <= 1e-6 and growth rate within 1% of FP64, momentum kept,
results reproducible bit for bit.

The scripts' quiet start is exactly uniform on the mixed path's fixed-point
lattice, so both precisions get the same seed, a 1e-8 mode-1 displacement of
the quiet start, and must grow at the same rate (within 1%) and at the
analytic one.  That seed is a few hundred charge quanta per cell, so the
quanta's rounding is a percent-level part of the tiny seed field: as for the
mixed midpoint, the quiet runs are held to the growth rate, and the
field-energy bar of 1e-6 is checked on a thermal workload (seeded x, +-1
beams plus N(0, 0.05^2), N=256, P=2^16) from the first step."""
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


def gamma_slope(W):
    """2 gamma / ln 10 of src/GaussianFixedPointQuietSimpson13.jl:19-20 for mode 1."""
    xk = 2 * np.pi / np.sqrt(W / 2)
    g = np.imag(np.sqrt(complex(xk ** 2 + 1 - np.sqrt(4 * xk ** 2 + 1)))) * np.sqrt(W / 2)
    return 2 * g / np.log(10)


@pytest.mark.parametrize("shape", ["gauss7", "area"])
def test_mixed_simpson_quiet_growth_rate(pg, shape):
    from picgolf import configs
    cf, x, v, _ = configs.quiet_simpson(shape)
    cm, _, _, _ = configs.quiet_simpson(shape, precision="mixed")
    cf.rtol = cm.rtol = 1e-8
    x = np.mod(x + 1e-8 * np.sin(2 * np.pi * x), 1.0)
    n = 3000
    Dm, Sm, *_ = run(pg, cm, x, v, n)
    Df, Sf, *_ = run(pg, cf, x, v, n)
    t = np.arange(1, n + 1) * cf.dt
    slopes = []
    for D in (Dm, Df):
        la = np.log10(D[:, 0])
        sel = (la > la[0] + 2) & (la < -3)
        assert sel.sum() > 200
        slopes.append(np.polyfit(t[sel], la[sel], 1)[0])
    rel = np.abs(Dm[:, 0] - Df[:, 0]) / Df[:, 0]
    print(f"Simpson {shape} mixed: growth slope {slopes[0]:.4f}, FP64 {slopes[1]:.4f}, analytic "
          f"{gamma_slope(cf.W):.4f}; field-energy rel err at the end {rel[-1]:.3e}; "
          f"sweeps {Sm.mean():.2f} vs {Sf.mean():.2f}; momentum {np.abs(Dm[:, 3]).max():.2e}")
    assert slopes[0] == pytest.approx(slopes[1], rel=0.01)
    assert slopes[1] == pytest.approx(gamma_slope(cf.W), rel=0.02)
    # momentum: the gather's FP32 weights and field against the exact quanta
    # of the deposit, a few 1e-9 once the instability has saturated
    assert np.abs(Dm[:, 3]).max() < 1e-8


@pytest.mark.parametrize("shape", ["gauss7", "area"])
def test_mixed_simpson_thermal_field_energy_and_determinism(pg, shape):
    from picgolf.core import Config
    N, P, W = 256, 2 ** 16, 400.0
    rng = np.random.default_rng(3)
    x = rng.random(P)
    v = np.where(np.arange(P) < P // 2, -1.0, 1.0) + 0.05 * rng.standard_normal(P)
    out = {}
    for prec in ("mixed", "fp64"):
        cfg = Config(N=N, P=P, W=W, dt=1 / (6 * N), shape=shape, scheme="simpson", precision=prec, rtol=1e-8)
        out[prec] = run(pg, cfg, x, v, 300)
    again = run(pg, Config(N=N, P=P, W=W, dt=1 / (6 * N), shape=shape, scheme="simpson", precision="mixed",
                           rtol=1e-8), x, v, 300)
    for a, b in zip(out["mixed"], again):
        assert np.array_equal(a, b)
    Dm, Df = out["mixed"][0], out["fp64"][0]
    rel = np.abs(Dm[:, 0] - Df[:, 0]) / Df[:, 0]
    print(f"Simpson {shape} thermal N={N} P={P}: field-energy rel err max {rel.max():.3e}; momentum drift "
          f"mixed {np.abs(Dm[:, 3] - Dm[0, 3]).max():.2e}")
    assert rel.max() <= 1e-6
    assert np.abs(Dm[:, 3] - Dm[0, 3]).max() < 1e-9
    # the final states agree to the velocity format
    vm, vf = out["mixed"][3], out["fp64"][3]
    assert np.max(np.abs(vm - vf)) < 1e-4


def test_mixed_simpson_halted_steps_match_full_slots(pg):
    """The mixed Simpson step through the halt-and-finish route equals a
    context that runs every sweep slot (bit for bit)."""
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_simpson("area", precision="mixed")
    N = cfg.N
    j = np.arange(N)
    E0 = 40.0 * np.sin(2 * np.pi * j / N)
    v2 = v + 0.5 * np.random.default_rng(12).standard_normal(v.size)
    outs = []
    for warm in (False, True):
        with pg.Context(cfg) as c:
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
