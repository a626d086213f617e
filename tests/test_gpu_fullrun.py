"""Whole runs at the configured lengths.

FP64: the quiet two-stream run (N=64, P=2048, 8192 steps, rtol=4eps) and the
NGP leapfrog run (N=128, P=8192, 1024 steps) bit for bit against the committed
golden fixtures of the CPU restatement (tests/golden/make_golden.py), through
both run paths (captured step graphs and the persistent kernel); the quiet
run's growth rate from the device trace against the analytic 2 pi/sqrt(3)
(src/GaussianFixedPointQuiet.jl:19-20); the scaled workload's whole
2000-step run at P=2^16 against its golden fixture, and at its full 2^24
particles bit for bit against the OpenMP restatement for two steps.

The persistent kernels against the oracle on the edge cases of their
layout: ragged particle counts, the smallest and largest grids, one
workgroup, several particle slots per lane group, the leapfrog Gaussian of
src/Gaussian.jl, and runs continued across calls, uploads and path changes.
"""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = float(np.finfo(np.float64).eps)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.fixture(scope="module")
def pg():
    import picgolf
    picgolf.load()
    return picgolf


def run_path(pg, cfg, x, v, nsteps, path, E=None, workgroups=0):
    with pg.Context(cfg) as c:
        got = c.set_run_path(path, workgroups)
        assert got[0] == ("graph" if path == "graph" else "persistent")
        c.upload(x, v, E)
        D, S = c.run(nsteps)
        xg, vg, Eg = c.download()
    return xg, vg, Eg, D[0], S[0]


def check_golden(name, got):
    g = np.load(os.path.join(GOLDEN, name))
    xg, vg, Eg, D, S = got
    ps = int(g["pstride"])
    assert np.array_equal(S, g["sweeps"].astype(np.int32))
    assert digest(S.astype(np.int32)) == str(g["S_sha256"])
    st = int(g["stride"])
    assert same(D[st - 1::st], g["D_every"])
    assert digest(D) == str(g["D_sha256"])
    assert same(xg[::ps], g["x"]) and same(vg[::ps], g["v"]) and same(Eg, g["E"])
    assert digest(xg) == str(g["x_sha256"]) and digest(vg) == str(g["v_sha256"]) and digest(Eg) == str(g["E_sha256"])


@pytest.mark.parametrize("path", ["graph", "persistent"])
def test_quiet_two_stream_full_run_matches_golden(pg, path):
    from picgolf import configs
    cfg, x, v, T = configs.quiet_two_stream()
    got = run_path(pg, cfg, x, v, T, path)
    check_golden("quiet_two_stream_8192.npz", got)
    # the growth the reference overplots: 2 gamma, gamma = 2 pi / sqrt(3)
    # (src/GaussianFixedPointQuiet.jl:19-20), in decades per unit time
    D = got[3]
    t = np.arange(1, T + 1) * cfg.dt
    la = np.log10(D[:, 0])
    sel = (t > 1.5) & (t < 6.5)
    slope = np.polyfit(t[sel], la[sel], 1)[0]
    analytic = 2 * (2 * np.pi / np.sqrt(3)) / np.log(10)
    print(f"{path}: quiet two-stream growth {slope:.4f} decades/time, analytic {analytic:.4f}")
    assert slope == pytest.approx(analytic, rel=0.01)
    assert np.max(np.abs(D[:, 3])) < 1e-14


@pytest.mark.parametrize("path", ["graph", "persistent"])
def test_ngp_two_stream_full_run_matches_golden(pg, path):
    from picgolf import configs
    cfg, x, v, T = configs.ngp_two_stream()
    check_golden("ngp_two_stream_1024.npz", run_path(pg, cfg, x, v, T, path))


def test_scaled_config_full_run_at_reduced_particles_matches_golden(pg):
    """The scaled workload's whole 2000-step FP64 run at P=2^16 (N=4096,
    Gaussian +-7 midpoint, rtol=1e-8) bit for bit against the golden fixture:
    every sweep count, the D trace, the final state and field."""
    from picgolf import configs
    cfg, x, v, T = configs.scaled_two_stream(precision="fp64", P=2 ** 16)
    with pg.Context(cfg) as c:
        assert c.run_path()[0] == "graph"
        c.upload(x, v)
        D, S = c.run(T)
        xg, vg, Eg = c.download()
    check_golden("scaled_two_stream_p65536_2000.npz", (xg, vg, Eg, D[0], S[0]))


def oracle_run(orc, cfg, x, v, nsteps, E=None, nthreads=8):
    shape = {"ngp": 0, "gauss6": 6, "gauss7": 7}[cfg.shape]
    scheme = 1 if cfg.scheme == "midpoint" else 0
    return orc.run(cfg.N, shape, scheme, cfg.W, cfg.w, cfg.dt, x, v, nsteps, rtol=cfg.rtol,
                   max_sweeps=cfg.max_sweeps, E=E, nthreads=nthreads)


def assert_run_equal(got, ref):
    xg, vg, Eg, D, S = got
    xo, vo, Eo, Do, So = ref
    assert np.array_equal(S, So)
    for a, b in zip((xg, vg, Eg, D), (xo, vo, Eo, Do)):
        assert same(a, b)


CASES = [
    # (N, P, shape, scheme, rtol, workgroups, steps): ragged P, extreme N,
    # one workgroup, several slots per lane group, the leapfrog Gaussian
    (8, 37, "gauss7", "midpoint", 1e-8, 0, 30),
    (256, 3000, "gauss7", "midpoint", 1e-10, 0, 12),
    (64, 1000, "gauss6", "midpoint", 4 * EPS, 8, 10),
    (64, 16384, "gauss7", "midpoint", 1e-8, 0, 6),
    (32, 2047, "gauss7", "midpoint", 1e-8, 16, 20),
    (128, 8192, "gauss6", "leapfrog", 1e-8, 0, 40),
    (16, 100, "gauss7", "leapfrog", 1e-8, 1, 25),
    (256, 8192, "ngp", "leapfrog", 1e-8, 0, 50),
    (8, 1, "ngp", "leapfrog", 1e-8, 0, 20),
    (64, 777, "ngp", "leapfrog", 1e-8, 0, 33),
]


@pytest.mark.parametrize("N,P,shape,scheme,rtol,wg,steps", CASES)
def test_persistent_against_oracle(pg, orc, N, P, shape, scheme, rtol, wg, steps):
    from picgolf.core import Config
    W = 300.0
    dt = 1 / (10 * N) if scheme == "leapfrog" else 1 / (6 * N)
    cfg = Config(N=N, P=P, W=W, w=W / P * N, dt=dt, shape=shape, scheme=scheme, rtol=rtol)
    rng = np.random.default_rng(N * 131 + P)
    x = rng.random(P)
    v = np.where(np.arange(P) < P // 2, -1.0, 1.0) + 0.05 * rng.standard_normal(P)
    print(f"persistent case N={N} P={P} {shape} {scheme} workgroups={wg}", flush=True)
    got = run_path(pg, cfg, x, v, steps, "persistent", workgroups=wg)
    assert_run_equal(got, oracle_run(orc, cfg, x, v, steps))


def test_persistent_continues_across_calls_uploads_and_paths(pg, orc):
    """Runs split over calls, an upload with a strong field (more sweeps)
    in between, and the run path switched mid-run: all bit for bit the
    oracle's continuous run."""
    from picgolf import configs
    cfg, x, v, _ = configs.quiet_two_stream()
    N = cfg.N
    with pg.Context(cfg) as c:
        c.upload(x, v)
        D1, S1 = c.run(13)
        c.set_run_path("graph")
        D2, S2 = c.run(9)
        c.set_run_path("persistent")
        D3, S3 = c.run(11)
        xm, vm, Em = c.download()
        ref = oracle_run(orc, cfg, x, v, 33)
        assert np.array_equal(np.concatenate([S1[0], S2[0], S3[0]]), ref[4])
        assert same(np.concatenate([D1[0], D2[0], D3[0]]), ref[3])
        assert same(xm, ref[0]) and same(vm, ref[1]) and same(Em, ref[2])
        j = np.arange(N)
        E0 = 40.0 * np.sin(2 * np.pi * j / N) + 15.0 * np.cos(6 * np.pi * j / N)
        v2 = v + 0.5 * np.random.default_rng(11).standard_normal(v.size)
        c.upload(x, v2, E0)
        D4, S4 = c.run(12)
        got = c.download()
    ref = oracle_run(orc, cfg, x, v2, 12, E=E0)
    assert S4.max() > S1.max()
    assert_run_equal((*got, D4[0], S4[0]), ref)


@pytest.mark.parametrize("path", ["graph", "persistent"])
def test_async_run_matches_golden_and_refuses_calls_in_flight(pg, path):
    """pic_run_async + pic_sync give the bits of pic_run; while the run is in
    flight every other call returns PIC_ESTATE and leaves the run alone."""
    from picgolf import configs
    from picgolf._lib import PicError
    cfg, x, v, T = configs.quiet_two_stream()
    with pg.Context(cfg) as c:
        c.set_run_path(path)
        c.upload(x, v)
        c.run_async(T)
        # 8192 steps take far longer than these calls
        for call in (c.download, lambda: c.run(1), lambda: c.upload(x, v), lambda: c.run_async(1)):
            with pytest.raises(PicError) as e:
                call()
            assert e.value.code == -5
        D, S = c.sync()
        assert c.sync() == (None, None)
        got = c.download()
    check_golden("quiet_two_stream_8192.npz", (*got, D[0], S[0]))


def test_run_path_selection(pg):
    from picgolf import configs
    from picgolf.core import Config
    cfg, _, _, _ = configs.quiet_two_stream()
    with pg.Context(cfg) as c:
        assert c.run_path() == ("persistent", 64)
        assert c.set_run_path("graph") == ("graph", 0)
        assert c.set_run_path("persistent", 128) == ("persistent", 128)
        with pytest.raises(Exception):
            c.set_run_path("persistent", 129)
    big = Config(N=4096, P=2 ** 16, W=400.0, dt=1 / (6 * 4096))
    with pg.Context(big) as c:
        assert c.run_path() == ("graph", 0)
        with pytest.raises(Exception):
            c.set_run_path("persistent")
    mixed = Config(N=64, P=2048, W=400.0, dt=1 / 384, precision="mixed")
    with pg.Context(mixed) as c:
        assert c.run_path()[0] == "graph"


def test_scaled_full_size_fp64_bitwise(pg, orc):
    """The scaled workload at full size (N=4096, P=2^24), two steps, bit for
    bit against the OpenMP restatement."""
    from picgolf import configs
    cfg, x, v, _ = configs.scaled_two_stream(precision="fp64")
    with pg.Context(cfg) as c:
        c.upload(x, v)
        D, S = c.run(2)
        xg, vg, Eg = c.download()
    ref = oracle_run(orc, cfg, x, v, 2, nthreads=16)
    assert_run_equal((xg, vg, Eg, D[0], S[0]), ref)
