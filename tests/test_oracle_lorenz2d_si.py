"""Pins the oracle of LorenzGaugeSemiImplicitField (src/PIC2D3V.jl:403-461,
warmup! :864-871, loop! :993-1100; oracle/lorenz2d.c orclg_si_*) against what
the mathematics and the module's expressions fix: the buffered deposits as
exact rational sums of the contributions on the grid 2^-q (per buffered
cell and per periodic cell, the stencil
layout of the OffsetArray axes -(b-1):NX+b, the BoundsError of a buffer too
small), warmup!'s three passes from the pinned deposits, single-mode closed
forms of the theta-weighted E and B and of the three source weights, a numpy
restatement of whole steps (floating-point deposits, numpy's FFT, isapprox
through numpy norms) with the same iteration counts, the stopping rule
(maxiters, rtol = 0), free streaming of uncharged particles before the first
update!, and split calls."""
from fractions import Fraction

import numpy as np
import pytest

from test_oracle_lorenz2d import mode, numpy_boris


def species(Lx, Ly, P=48, seed=5, shape="bspline2", weight=0.01):
    rng = np.random.default_rng(seed)
    out = []
    for q, m, vth in ((-1.0, 1.0, 0.3), (1.0, 32.0, 0.05)):
        xyv = np.vstack([rng.uniform(0, Lx, P), rng.uniform(0, Ly, P), vth * rng.standard_normal((3, P))])
        out.append(dict(shape=shape, charge=q, mass=m, weight=weight, xyv=xyv))
    return out


@pytest.mark.parametrize("shape", ["ngp", "area", "bspline2", "bspline5"])
def test_buffered_deposit_exact(orc, shape):
    NX, NY, Lx, Ly, b = 8, 4, 2.0, 1.0, 3
    sp = species(Lx, Ly, P=12, seed=1, shape=shape)
    sp[0]["xyv"][0, :2] = [Lx, 1e-9]                  # next to both ends: ghost cells
    sp[0]["xyv"][1, 2:4] = [Ly, 1e-9]
    rc, grid, per = orc.lg_si_deposit(NX, NY, Lx, Ly, sp, b)
    assert rc == 0
    dV = (Lx / NX) * (Ly / NY)
    NXe, NYe = NX + 2 * b, NY + 2 * b
    # each contribution on the grid 2^-q (ties to even), q from the bound
    # sum_s P_s |qw_s| (times max |v_c| for the currents)
    P = [s["xyv"].shape[1] for s in sp]
    qws = [s["charge"] * s["weight"] / dV for s in sp]
    qs = [orc.lg_frac_bits(P, qws)] + [orc.lg_frac_bits(P, qws, [np.abs(s["xyv"][2 + c]).max() for s in sp])
                                       for c in range(3)]
    acc = [[Fraction(0)] * (NXe * NYe) for _ in range(4)]
    for s in sp:
        x, y, vx, vy, vz = s["xyv"]
        qw = s["charge"] * s["weight"] / dV
        for p in range(x.size):
            cx, wx = orc.weights23(shape, x[p], NX / Lx)
            cy, wy = orc.weights23(shape, y[p], NY / Ly)
            for c, w in enumerate((qw, vx[p] * qw, vy[p] * qw, vz[p] * qw)):
                for j, bb in zip(cy, wy):
                    for i, a in zip(cx, wx):
                        # Julia index i in -(b-1):NX+b is row i + b - 1
                        t = Fraction(round(Fraction(a * bb * w) * 2 ** qs[c]), 2 ** qs[c])
                        acc[c][(i + b - 1) + NXe * (j + b - 1)] += t
    for c in range(4):
        assert np.array_equal(grid[:, :, c].ravel(), np.array([float(a) for a in acc[c]]))
        wrapped = [Fraction(0)] * (NX * NY)
        for e, a in enumerate(acc[c]):
            i, j = e % NXe - (b - 1), e // NXe - (b - 1)
            wrapped[(i - 1) % NX + NX * ((j - 1) % NY)] += a
        assert np.array_equal(per[c].real.ravel(), np.array([float(a) for a in wrapped]))
        assert not np.any(per[c].imag)
    if shape != "ngp":
        assert np.any(grid[:b]) or np.any(grid[:, :b]) or np.any(grid[-b:]) or np.any(grid[:, -b:])
    # the currents alone leave the charge component zero
    rc, g2, p2 = orc.lg_si_deposit(NX, NY, Lx, Ly, sp, b, currents_only=True)
    assert rc == 0 and not np.any(g2[:, :, 0]) and not np.any(p2[0])
    assert np.array_equal(g2[:, :, 1:], grid[:, :, 1:]) and np.array_equal(p2[1:], per[1:])


def test_buffer_too_small_is_an_error(orc):
    NX, NY, Lx, Ly = 8, 8, 1.0, 1.0
    sp = species(Lx, Ly, P=8, seed=2, shape="bspline2")
    sp[0]["xyv"][0, 0] = Lx                          # stencil reaches Julia index NX + 2
    assert orc.lg_si_deposit(NX, NY, Lx, Ly, sp, 1)[0] == -2
    assert orc.lg_si_deposit(NX, NY, Lx, Ly, sp, 2)[0] == 0
    with pytest.raises(ValueError):
        orc.lg_si_warmup(NX, NY, Lx, Ly, 0.01, sp, 1)


def test_warmup_passes(orc):
    NX, NY, Lx, Ly, dt, b = 16, 8, 1.0, 2.0, 0.02, 3
    sp = species(Lx, Ly, P=40, seed=3)
    sp1, F, grid = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp, b)
    u = orc.unimod23
    cur = [dict(s, xyv=s["xyv"].copy()) for s in sp]
    outs = []
    for h in (-dt, dt, dt):
        for s in cur:
            x, y, vx, vy, _ = s["xyv"]
            s["xyv"][0] = [u(a + v * h, Lx) for a, v in zip(x, vx)]
            s["xyv"][1] = [u(a + v * h, Ly) for a, v in zip(y, vy)]
        outs.append(orc.lg_si_deposit(NX, NY, Lx, Ly, cur, b))
        for s in cur:
            x, y, vx, vy, _ = s["xyv"]
            s["xyv"][0] = [u(a - v * h, Lx) for a, v in zip(x, vx)]
            s["xyv"][1] = [u(a - v * h, Ly) for a, v in zip(y, vy)]
    for s in cur:
        x, y, vx, vy, _ = s["xyv"]
        s["xyv"][0] = [u(a + v * -dt, Lx) for a, v in zip(x, vx)]
        s["xyv"][1] = [u(a + v * -dt, Ly) for a, v in zip(y, vy)]
    for a, c in zip(sp1, cur):
        assert np.array_equal(a["xyv"], c["xyv"])
    assert np.array_equal(F[18:22], outs[0][2]) and np.array_equal(F[8:12], outs[1][2])
    assert np.array_equal(grid, outs[2][1])
    assert not np.any(F[:8]) and not np.any(F[12:18])
    assert np.any(F[8]) and np.any(F[18])


def test_solve_single_mode_closed_forms(orc):
    NX, NY, Lx, Ly, dt = 16, 8, 2.0, 0.5, 0.01
    z = np.zeros((4, NY, NX), dtype=complex)
    for th in (0.0, 0.5, 1.0):
        for mx, my in ((1, 0), (3, 0), (0, 2), (2, 1)):
            kx0, ky0 = 2 * np.pi * mx / Lx, 2 * np.pi * my / Ly
            k2 = kx0 ** 2 + ky0 ** 2
            r = (2 - dt ** 2 * k2 * (1 - th)) / (1 + dt ** 2 * k2 * th / 2)
            w = th / 2 * r + (1 - th)            # theta/2 (x+ + x-) + (1 - theta) x0, x- = 0
            c, s = mode(NX, NY, Lx, Ly, mx, my), mode(NX, NY, Lx, Ly, mx, my, "sin")
            F = np.zeros((22, NY, NX), dtype=complex)
            F[0] = np.fft.fft2(c)
            F1, _, plus = orc.lg_si_solve(F, z, dt, th, 0.3, Lx, Ly)
            assert np.allclose(plus[0], r * F[0], atol=1e-12 * NX * NY)
            assert np.allclose(F1[12].real, w * kx0 * s, atol=1e-11)
            assert np.allclose(F1[13].real, w * ky0 * s, atol=1e-11)
            # Az0 = cos: Ez = -(Az+ - Az-)/2dt, Bx = d(Azc)/dy, By = -d(Azc)/dx
            F = np.zeros((22, NY, NX), dtype=complex)
            F[4] = np.fft.fft2(c)
            F1, _, _ = orc.lg_si_solve(F, z, dt, th, 0.3, Lx, Ly)
            assert np.allclose(F1[14].real, -r / (2 * dt) * c, atol=1e-9)
            assert np.allclose(F1[15].real, -w * ky0 * s, atol=1e-11)
            assert np.allclose(F1[16].real, w * kx0 * s, atol=1e-11)
            # Ax- = cos: Ex = +Ax-/2dt, Bz = -d/dy(theta/2 Ax-) and x+ = -x-
            F = np.zeros((22, NY, NX), dtype=complex)
            F[5] = np.fft.fft2(c)
            F1, _, plus = orc.lg_si_solve(F, z, dt, th, 0.3, Lx, Ly)
            assert np.array_equal(plus[1], -F[5])
            assert np.allclose(F1[12].real, (1 - (-1)) / (2 * dt) * c, atol=1e-9)
            assert np.allclose(F1[17].real, -(th / 2 * (-1 + 1)) * ky0 * s, atol=1e-11)
    # the three source weights: s- and s0 enter as given, s+ through its transform with [1,1] = 0
    rng = np.random.default_rng(4)
    S = rng.standard_normal((NY, NX))
    _, _, k2 = orc.lg_wavenumbers(NX, NY, Lx, Ly)
    th, ts = 1.0, 0.05
    den = 1 + dt ** 2 * k2 * th / 2
    for slot, weight in ((18, ts / 2), (8, 1 - ts)):
        F = np.zeros((22, NY, NX), dtype=complex)
        F[slot] = S
        _, _, plus = orc.lg_si_solve(F, z, dt, th, ts, Lx, Ly)
        assert np.allclose(plus[0], dt ** 2 * weight * S / den, rtol=1e-15, atol=0)
    src = z.copy()
    src[1] = S
    F1, src1, plus = orc.lg_si_solve(np.zeros((22, NY, NX), dtype=complex), src, dt, th, ts, Lx, Ly)
    spec = np.fft.fft2(S)
    spec[0, 0] = 0
    assert np.allclose(src1[1], spec, atol=1e-12) and src1[1][0, 0] == 0
    assert np.allclose(plus[1], dt ** 2 * ts / 2 * spec / den, atol=1e-16)


def wrap(NX, NY, b, a):
    out = np.zeros((NY, NX))
    NYe, NXe = a.shape
    for jj in range(NYe):
        for ii in range(NXe):
            out[(jj - b) % NY, (ii - b) % NX] += a[jj, ii]
    return out


def numpy_si_steps(orc, NX, NY, Lx, Ly, dt, B0, th, ts, rtol, maxiters, b, species_, F, grid, EB, T):
    """loop! in numpy: floating-point deposits into the buffered grid, numpy's
    FFT and isapprox through numpy norms."""
    kx = (2 * np.pi / Lx) * np.concatenate([np.arange(NX // 2), np.arange(-NX // 2, 0)])
    ky = (2 * np.pi / Ly) * np.concatenate([np.arange(NY // 2), np.arange(-NY // 2, 0)])
    KX, KY = kx[None, :], ky[:, None]
    k2 = KX ** 2 + KY ** 2
    dV = (Lx / NX) * (Ly / NY)
    F, grid, EB = F.copy(), grid.copy(), EB.copy()
    sp = [dict(s, xyv=s["xyv"].copy()) for s in species_]
    num = 2 - dt ** 2 * k2 * (1 - th)
    den = 1 + dt ** 2 * k2 * th / 2
    iters = []
    for _ in range(T):
        k, conv = 0, False
        while not (k > 0 and (k > maxiters or conv)):
            k += 1
            new = np.zeros_like(grid)
            work = []
            for s in sp:
                x, y, vx, vy, vz = s["xyv"]
                qw = s["charge"] * s["weight"] / dV
                q_m = s["charge"] / s["mass"]
                wk = np.zeros_like(s["xyv"])
                for p in range(x.size):
                    g = np.array([orc.gather23(EB[c], EB[c + 1], x[p], y[p], s["shape"], Lx, Ly)
                                  for c in (0, 2, 4)]).ravel()
                    wk[0, p] = orc.unimod23(x[p] + vx[p] * dt, Lx)
                    wk[1, p] = orc.unimod23(y[p] + vy[p] * dt, Ly)
                    wk[2:, p] = numpy_boris(dt, q_m, g[:3], g[3:], np.array([vx[p], vy[p], vz[p]]))
                    cx, wx = orc.weights23(s["shape"], wk[0, p], NX / Lx)
                    cy, wy = orc.weights23(s["shape"], wk[1, p], NY / Ly)
                    for j, bb in zip(cy, wy):
                        for i, a in zip(cx, wx):
                            new[j + b - 1, i + b - 1, 1:] += a * bb * wk[2:, p] * qw
                work.append(wk)
            d = np.linalg.norm(grid - new)
            conv = d <= rtol * max(np.linalg.norm(grid), np.linalg.norm(new))
            grid = new
            src = [np.zeros((NY, NX), dtype=complex)] + [np.fft.fft2(wrap(NX, NY, b, new[:, :, c]))
                                                         for c in (1, 2, 3)]
            for a in src:
                a[0, 0] = 0
            plus = []
            for a, (x0, xm, s0, sm) in enumerate(((0, 1, 8, 18), (2, 5, 9, 19), (3, 6, 10, 20), (4, 7, 11, 21))):
                sv = ts / 2 * F[sm] + (1 - ts) * F[s0] + ts / 2 * src[a]
                plus.append((num * F[x0] + dt ** 2 * sv) / den - F[xm])
            c = [th / 2 * (plus[a] + F[xm]) + (1 - th) * F[x0] for a, (x0, xm) in enumerate(((0, 1), (2, 5), (3, 6),
                                                                                             (4, 7)))]
            E = [-1j * KX * c[0] - (plus[1] - F[5]) / (2 * dt), -1j * KY * c[0] - (plus[2] - F[6]) / (2 * dt),
                 -(plus[3] - F[7]) / (2 * dt)]
            B = [1j * KY * c[3], -1j * KX * c[3], 1j * KX * c[2] - 1j * KY * c[1]]
            for q in range(3):
                F[12 + q] = np.fft.ifft2(E[q])
                F[15 + q] = np.fft.ifft2(B[q])
            EB = np.array([F[12 + q].real + (B0[q - 3] if q >= 3 else 0.0) for q in range(6)])
        iters.append(k)
        for s, wk in zip(sp, work):
            s["xyv"] = wk
        F[1], F[5], F[6], F[7] = F[0], F[2], F[3], F[4]
        F[0], F[2], F[3], F[4] = plus
    return sp, F, grid, EB, np.array(iters)


@pytest.mark.parametrize("th,ts", [(1.0, 0.05), (0.5, 0.0), (1.0, 1.0)])
def test_steps_against_numpy(orc, th, ts):
    NX, NY, Lx, Ly, dt, b, T = 16, 16, 1.0, 2.0, 0.02, 3, 3
    B0 = np.array([0.6, 0.0, 0.2])
    sp, F, grid = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, species(Lx, Ly, P=40, seed=8), b)
    EB = np.zeros((6, NY, NX))
    rtol = float(np.sqrt(np.finfo(float).eps))
    o = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp, F, grid, EB, T, theta=th, theta_source=ts, rtol=rtol, buffer=b,
                      diagnose=False)
    n = numpy_si_steps(orc, NX, NY, Lx, Ly, dt, B0, th, ts, rtol, 10, b, sp, F, grid, EB, T)
    assert np.array_equal(o[6], n[4]) and np.all(o[6] > 1) and np.all(o[6] <= 11)
    for a, c in zip(o[0], n[0]):
        assert np.allclose(a["xyv"], c["xyv"], rtol=1e-10, atol=1e-13)
    scale = np.abs(n[1]).max(axis=(1, 2), keepdims=True) + 1e-300
    assert np.all(np.abs(o[1] - n[1]) <= 1e-9 * scale)
    assert np.allclose(o[2], n[2], rtol=1e-12, atol=1e-12 * np.abs(n[2]).max())
    assert np.allclose(o[3], n[3], rtol=1e-9, atol=1e-12)
    assert np.abs(o[1][12:18]).max() > 0 and np.abs(o[2][:, :, 1:]).max() > 0


def test_stopping_rule_and_free_streaming(orc):
    NX, NY, Lx, Ly, dt, b = 16, 8, 1.0, 1.0, 0.02, 3
    B0 = np.array([0.0, 0.0, 2.0])
    sp, F, grid = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, species(Lx, Ly, P=24, seed=9), b)
    EB = np.zeros((6, NY, NX))
    # rtol = 0: only an exact repeat of the grid stops the loop before maxiters
    for maxiters in (0, 1, 2):
        o = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp, F, grid, EB, 2, rtol=0.0, maxiters=maxiters, buffer=b,
                          diagnose=False)
        assert np.array_equal(o[6], [maxiters + 1] * 2)
    o = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp, F, grid, EB, 1, rtol=0.0, maxiters=50, buffer=b, diagnose=False)
    k = int(o[6][0])
    assert 3 <= k < 51
    # ... and the grid of iteration k is that of iteration k - 1 exactly
    p = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp, F, grid, EB, 1, rtol=0.0, maxiters=k - 2, buffer=b,
                      diagnose=False)
    assert np.array_equal(p[2], o[2])
    # uncharged particles: the grids stay zero (converged after one
    # iteration), the first step streams freely (EBxyz is still zero), the
    # second turns about B0
    sp0 = species(Lx, Ly, P=24, seed=10, weight=0.0)
    sp1, F0, grid0 = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, sp0, b)
    assert not np.any(grid0) and not np.any(F0)
    o = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp1, F0, grid0, EB, 1, buffer=b, diagnose=False)
    assert np.array_equal(o[6], [1])
    for a, c in zip(o[0], sp1):
        x, y, vx, vy, vz = c["xyv"]
        assert np.array_equal(a["xyv"][0], [orc.unimod23(p + v * dt, Lx) for p, v in zip(x, vx)])
        assert np.array_equal(a["xyv"][2:], c["xyv"][2:])
    o2 = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, o[0], o[1], o[2], o[3], 1, buffer=b, diagnose=False)
    for a, c in zip(o2[0], o[0]):
        q_m = c["charge"] / c["mass"]
        ph = 2 * np.arctan(q_m * 2.0 * dt / 2)
        z0 = c["xyv"][2] + 1j * c["xyv"][3]
        assert np.allclose(a["xyv"][2] + 1j * a["xyv"][3], z0 * np.exp(-1j * ph), rtol=0, atol=1e-14)


def test_diagnostics_and_two_calls_equal_one(orc):
    NX, NY, Lx, Ly, dt, b, T = 16, 8, 1.0, 1.0, 0.02, 3, 4
    B0 = np.array([0.3, 0.0, 0.4])
    sp, F, grid = orc.lg_si_warmup(NX, NY, Lx, Ly, dt, species(Lx, Ly, P=32, seed=11), b)
    EB = np.zeros((6, NY, NX))
    o = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp, F, grid, EB, T, theta=1.0, theta_source=0.05, buffer=b,
                      ngskip=2)
    a = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp, F, grid, EB, 1, theta=1.0, theta_source=0.05, buffer=b,
                      ngskip=2, nbins=T)
    c = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, a[0], a[1], a[2], a[3], T - 1, theta=1.0, theta_source=0.05,
                      buffer=b, ngskip=2, t0=1, bins=a[5])
    for k in (1, 2, 3, 5):
        assert np.array_equal(c[k], o[k])
    assert np.array_equal(np.concatenate([a[4], c[4]]), o[4])
    assert np.array_equal(np.concatenate([a[6], c[6]]), o[6])
    # diagnose! transforms rho0, J0 (the constant warmup! sources) in place
    d = orc.lg_si_run(NX, NY, Lx, Ly, dt, B0, sp, F, grid, EB, 1, theta=1.0, theta_source=0.05, buffer=b,
                      diagnose=False)
    for k in (8, 9, 10, 11):
        assert np.array_equal(a[1][k], orc.lg_ifft2(d[1][k]))
    for k in (18, 19, 20, 21):
        assert np.array_equal(a[1][k], F[k]) and np.array_equal(d[1][k], F[k])
    # the row: the field energy of the ghosted EB and the kinetic energy
    row = orc.lg_row(a[1][:18], a[3], a[0], b)
    assert np.array_equal(row, a[4][0])
