"""Pins the oracle of the Lorenz-gauge electromagnetic 2D3V path
(oracle/lorenz2d.c, src/PIC2D3V.jl LorenzGaugeStaggeredField as src/2D3V.jl
drives it) against what the mathematics and the module's expressions fix:
the FFTHelper wavenumbers, closed forms of ElectromagneticBoris, the gather
against the pinned electrostatic gather, single-mode closed forms of
E = -grad phi - dA/dt and B = curl A in an Lx x Ly box (numpy's FFT), the
per-mode recurrence of the wave update (stable for theta >= 1/2, not for the
explicit scheme past k dt = 2), the source used as the spectrum, diagnose!'s
in-place inverse transforms, the exact warmup currents (rational sums),
a numpy restatement of whole steps with diagnostics, and split calls."""
from fractions import Fraction

import numpy as np
import pytest


def test_wavenumbers(orc):
    NX, NY, Lx, Ly = 8, 4, 2.0, 0.5
    kx, ky, k2 = orc.lg_wavenumbers(NX, NY, Lx, Ly)
    assert np.array_equal(kx, (2 * np.pi / Lx) * np.array([0, 1, 2, 3, -4, -3, -2, -1]))
    assert np.array_equal(ky, (2 * np.pi / Ly) * np.array([0, 1, -2, -1]))
    assert np.array_equal(k2, kx[None, :] ** 2 + ky[:, None] ** 2)


def test_boris_closed_forms(orc):
    dt = 0.1
    rng = np.random.default_rng(0)
    for q_m in (1.0, -1.0, 1 / 32):
        # E = 0: |v| kept; the turn about B by 2 atan(|t|), clockwise for q_m > 0
        b = 1.7
        v = orc.lg_boris(dt, q_m, np.zeros(3), np.array([0.0, 0.0, b]), np.array([1.0, 0.0, 0.3]))
        phi = 2 * np.arctan(q_m * b * dt / 2)
        assert np.allclose(v, [np.cos(phi), -np.sin(phi), 0.3], rtol=0, atol=1e-15)
        u = rng.standard_normal(3)
        Bv = rng.standard_normal(3)
        v = orc.lg_boris(dt, q_m, np.zeros(3), Bv, u)
        assert abs(np.linalg.norm(v) - np.linalg.norm(u)) < 1e-15 * 4
        assert abs(v @ Bv - u @ Bv) < 1e-14
        # B = 0: a plain kick
        E = rng.standard_normal(3)
        v = orc.lg_boris(dt, q_m, E, np.zeros(3), u)
        assert np.allclose(v, u + q_m * E * dt, rtol=0, atol=1e-15)
    # E x B drift of the averaged velocity
    E, B = np.array([0.0, 0.02, 0.0]), np.array([0.0, 0.0, 1.0])
    v, vs = np.array([0.3, 0.0, 0.0]), []
    for _ in range(20000):
        v = orc.lg_boris(0.01, 1.0, E, B, v)
        vs.append(v)
    drift = np.mean(vs[-12566:], axis=0)          # about 2 gyro-periods
    assert np.allclose(drift, np.cross(E, B) / (B @ B), atol=2e-3)


def test_gather_matches_the_pinned_electrostatic_gather(orc):
    NX, NY, Lx, Ly = 16, 8, 2.0, 1.0
    rng = np.random.default_rng(1)
    EB = rng.standard_normal((6, NY, NX))
    for shape in ("ngp", "area", "bspline1", "bspline2", "bspline3"):
        for _ in range(20):
            x, y = rng.uniform(0, Lx), rng.uniform(0, Ly)
            g = orc.lg_gather(EB, x, y, shape, Lx, Ly)
            for c in (0, 2, 4):
                a, b = orc.gather23(EB[c], EB[c + 1], x, y, shape, Lx, Ly)
                assert g[c] == a and g[c + 1] == b


def mode(NX, NY, Lx, Ly, mx=0, my=0, kind="cos"):
    x = np.arange(NX) * Lx / NX
    y = np.arange(NY) * Ly / NY
    X, Y = np.meshgrid(x, y)
    th = 2 * np.pi * (mx * X / Lx + my * Y / Ly)
    return np.cos(th) if kind == "cos" else np.sin(th)


def test_solve_single_mode_closed_forms(orc):
    NX, NY, Lx, Ly, dt = 16, 8, 2.0, 0.5, 0.01
    for theta in (0.0, 0.5, 1.0):
        for mx, my in ((1, 0), (3, 0), (0, 2)):
            kx0, ky0 = 2 * np.pi * mx / Lx, 2 * np.pi * my / Ly
            k2 = kx0 ** 2 + ky0 ** 2
            r = (2 - dt ** 2 * k2 * (1 - theta)) / (1 + dt ** 2 * k2 * theta / 2)
            c, s = mode(NX, NY, Lx, Ly, mx, my), mode(NX, NY, Lx, Ly, mx, my, "sin")
            F = np.zeros((18, NY, NX), dtype=complex)
            F[0] = np.fft.fft2(c)                   # phi0: phi+ = r phi0
            F1, plus = orc.lg_solve(F, dt, theta, Lx, Ly)
            assert np.allclose(plus[0], r * F[0], atol=1e-12 * NX * NY)
            assert np.allclose(F1[12].real, r * kx0 * s, atol=1e-11)     # Ex = -d(phi+)/dx
            assert np.allclose(F1[13].real, r * ky0 * s, atol=1e-11)
            assert np.allclose(F1[12].imag, 0, atol=1e-12)
            # A: Az0 = cos -> Ez = -(r - 1)/dt cos, Bx = dAz/dy, By = -dAz/dx
            F = np.zeros((18, NY, NX), dtype=complex)
            F[4] = np.fft.fft2(c)
            F1, _ = orc.lg_solve(F, dt, theta, Lx, Ly)
            assert np.allclose(F1[14].real, -(r - 1) / dt * c, atol=1e-9)
            assert np.allclose(F1[15].real, -(1 + r) / 2 * ky0 * s, atol=1e-11)
            assert np.allclose(F1[16].real, (1 + r) / 2 * kx0 * s, atol=1e-11)
            # Bz = dAy/dx - dAx/dy
            F = np.zeros((18, NY, NX), dtype=complex)
            F[3] = np.fft.fft2(c)
            F[2] = np.fft.fft2(2 * c)
            F1, _ = orc.lg_solve(F, dt, theta, Lx, Ly)
            assert np.allclose(F1[17].real, (1 + r) / 2 * (-kx0 * s + 2 * ky0 * s), atol=1e-11)
            assert np.allclose(F1[12].real, -(r - 1) / dt * 2 * c, atol=1e-9)
    # the source enters as given (the spectrum the solve sees is the real-space
    # array when no transform ran): phi+ = dt^2 rho / den, A+ = dt^2 J / den
    rng = np.random.default_rng(2)
    F = np.zeros((18, NY, NX), dtype=complex)
    F[8:12] = rng.standard_normal((4, NY, NX))
    _, _, k2 = orc.lg_wavenumbers(NX, NY, Lx, Ly)
    F1, plus = orc.lg_solve(F, dt, 1.0, Lx, Ly)
    den = 1 + dt ** 2 * k2 / 2
    assert np.allclose(plus, dt ** 2 * F[8:12] / den, rtol=1e-15, atol=0)
    # phi- enters with a minus sign
    F = np.zeros((18, NY, NX), dtype=complex)
    F[1] = rng.standard_normal((NY, NX))
    _, plus = orc.lg_solve(F, dt, 1.0, Lx, Ly)
    assert np.array_equal(plus[0], -F[1])


def inert_species(Lx, Ly, P=4, seed=3):
    rng = np.random.default_rng(seed)
    xyv = np.vstack([rng.uniform(0, Lx, P), rng.uniform(0, Ly, P), rng.standard_normal((3, P))])
    return [dict(shape="bspline2", charge=1.0, mass=1.0, weight=0.0, xyv=xyv)]


@pytest.mark.parametrize("theta,dt,stable", [(1.0, 0.05, True), (0.5, 0.05, True), (0.0, 0.005, True),
                                             (0.0, 0.05, False)])
def test_vacuum_recurrence(orc, theta, dt, stable):
    """No charge (weight 0), no diagnostics: each mode follows
    x+ = r x0 - x-, r = num/den: |lambda| = 1 while |r| < 2 (always for
    theta >= 1/2), growth for the explicit scheme once dt k > 2."""
    NX, NY, Lx, Ly, T = 16, 8, 1.0, 1.0, 60
    m = 7
    k2 = (2 * np.pi * m / Lx) ** 2
    r = (2 - dt ** 2 * k2 * (1 - theta)) / (1 + dt ** 2 * k2 * theta / 2)
    F = np.zeros((18, NY, NX), dtype=complex)
    F[0, 0, m] = 1.0
    F[1, 0, m] = 0.5
    sp = inert_species(Lx, Ly)
    _, F1, rows, _ = orc.lg_run(NX, NY, Lx, Ly, dt, [0, 0, 0], sp, F, T, theta=theta, diagnose=False)
    M = np.array([[r, -1.0], [1.0, 0.0]])
    s = np.linalg.matrix_power(M, T) @ np.array([1.0, 0.5])
    assert np.allclose([F1[0, 0, m], F1[1, 0, m]], s, rtol=1e-9, atol=1e-12)
    grows = abs(r) > 2
    assert grows == (not stable)
    if stable:
        assert np.max(np.abs(np.linalg.eigvals(M))) == pytest.approx(1.0, abs=1e-12)
    assert rows.shape == (0, 11)


def random_species(Lx, Ly, P=64, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for q, m, vth in ((-1.0, 1.0, 0.3), (1.0, 32.0, 0.05)):
        xyv = np.vstack([rng.uniform(0, Lx, P), rng.uniform(0, Ly, P), vth * rng.standard_normal((3, P))])
        out.append(dict(shape="bspline2", charge=q, mass=m, weight=0.01, xyv=xyv))
    return out


def test_diagnose_transforms_the_state_in_place(orc):
    NX, NY, Lx, Ly, dt = 16, 16, 1.0, 1.0, 0.01
    sp = random_species(Lx, Ly)
    rng = np.random.default_rng(6)
    F = rng.standard_normal((18, NY, NX)) + 1j * rng.standard_normal((18, NY, NX))
    B0 = [0.5, 0.0, 0.0]
    a = orc.lg_run(NX, NY, Lx, Ly, dt, B0, sp, F, 1, diagnose=False)
    b = orc.lg_run(NX, NY, Lx, Ly, dt, B0, sp, F, 1, diagnose=True)
    for i in range(2):
        assert np.array_equal(a[0][i]["xyv"], b[0][i]["xyv"])
    for k in (2, 3, 4, 0, 8, 9, 10, 11):                 # Ax0 Ay0 Az0 phi0 rho0 Jx0 Jy0 Jz0
        assert np.array_equal(b[1][k], orc.lg_ifft2(a[1][k]))
        assert np.allclose(b[1][k], np.fft.ifft2(a[1][k]), rtol=1e-12, atol=1e-15 * np.abs(a[1][k]).max())
    for k in (1, 5, 6, 7, 12, 13, 14, 15, 16, 17):       # phi- A- and the fields: untouched
        assert np.array_equal(a[1][k], b[1][k])
    # rho0 and J0 of loop! are real (imaginary parts exactly 0)
    assert not np.any(a[1][8:12].imag)


def test_warmup_exact_currents(orc):
    NX, NY, Lx, Ly, dt = 8, 8, 2.0, 1.0, 0.05
    sp = random_species(Lx, Ly, P=24, seed=7)
    sp1, F = orc.lg_warmup(NX, NY, Lx, Ly, dt, sp)
    assert not np.any(F[8]) and not np.any(F[9:12].imag)
    assert not np.any(F[:8]) and not np.any(F[12:])
    dV = (Lx / NX) * (Ly / NY)
    acc = [[Fraction(0)] * (NX * NY) for _ in range(3)]
    for s0, s1 in zip(sp, sp1):
        x, y, vx, vy, vz = s0["xyv"]
        P = x.size
        qw = s0["charge"] * s0["weight"] / dV
        for p in range(P):
            xb = orc.unimod23(orc.unimod23(x[p] + vx[p] * (-(dt / 2)), Lx) + vx[p] * (dt / 2), Lx)
            yb = orc.unimod23(orc.unimod23(y[p] + vy[p] * (-(dt / 2)), Ly) + vy[p] * (dt / 2), Ly)
            assert s1["xyv"][0, p] == xb and s1["xyv"][1, p] == yb
            cx, wx = orc.weights23("bspline2", xb, NX / Lx)
            cy, wy = orc.weights23("bspline2", yb, NY / Ly)
            for c, v in enumerate((vx[p], vy[p], vz[p])):
                w = v * qw
                for j, b in zip(cy, wy):
                    for i, a in zip(cx, wx):
                        acc[c][(i - 1) % NX + NX * ((j - 1) % NY)] += Fraction(a * b * w)
        assert np.array_equal(s1["xyv"][2:], s0["xyv"][2:])
    for c in range(3):
        assert np.array_equal(F[9 + c].real.ravel(), np.array([float(a) for a in acc[c]]))


def numpy_boris(dt, q_m, E, B, v):
    th = dt / 2 * q_m
    t = B * th
    s = 2 / (1 + t @ t)
    e = E * th
    m = v + e
    return m + np.cross(m + np.cross(m, t), t) * s + e


def numpy_steps(orc, NX, NY, Lx, Ly, dt, B0, theta, buffer, ngskip, species, F, T):
    """loop! and diagnose! in numpy: numpy's FFT, floating-point deposits,
    stencils and unimod from the pinned electrostatic oracle."""
    kx, ky, k2 = (2 * np.pi / Lx) * np.concatenate([np.arange(NX // 2), np.arange(-NX // 2, 0)]), \
        (2 * np.pi / Ly) * np.concatenate([np.arange(NY // 2), np.arange(-NY // 2, 0)]), None
    KX, KY = kx[None, :], ky[:, None]
    k2 = KX ** 2 + KY ** 2
    dV = (Lx / NX) * (Ly / NY)
    F = F.copy()
    sp = [dict(s, xyv=s["xyv"].copy()) for s in species]
    rows, bins = [], np.zeros((14, T, NY // ngskip, NX // ngskip))
    for t in range(T):
        num = 2 - dt ** 2 * k2 * (1 - theta)
        den = 1 + dt ** 2 * k2 * theta / 2
        plus = [(num * F[x0] + dt ** 2 * F[s]) / den - F[xm] for x0, xm, s in ((0, 1, 8), (2, 5, 9), (3, 6, 10),
                                                                               (4, 7, 11))]
        ph, ax, ay, az = plus
        E = [-1j * KX * ph - (ax - F[2]) / dt, -1j * KY * ph - (ay - F[3]) / dt, -(az - F[4]) / dt]
        B = [1j * KY * (az + F[4]) / 2, -1j * KX * (az + F[4]) / 2,
             1j * KX * (ay + F[3]) / 2 - 1j * KY * (ax + F[2]) / 2]
        for c in range(3):
            F[12 + c] = np.fft.ifft2(E[c])
            F[15 + c] = np.fft.ifft2(B[c])
        EB = np.array([F[12 + c].real + (B0[c - 3] if c >= 3 else 0.0) for c in range(6)])
        rho, J = np.zeros((NY, NX)), np.zeros((3, NY, NX))
        for s in sp:
            x, y, vx, vy, vz = s["xyv"]
            qw = s["charge"] * s["weight"] / dV
            q_m = s["charge"] / s["mass"]
            for p in range(x.size):
                g = np.array([orc.gather23(EB[c], EB[c + 1], x[p], y[p], s["shape"], Lx, Ly) for c in (0, 2, 4)]).ravel()
                v = numpy_boris(dt, q_m, g[:3], g[3:], np.array([vx[p], vy[p], vz[p]]))
                vx[p], vy[p], vz[p] = v
                for half in range(2):
                    x[p] = orc.unimod23(x[p] + vx[p] * dt / 2, Lx)
                    y[p] = orc.unimod23(y[p] + vy[p] * dt / 2, Ly)
                    cx, wx = orc.weights23(s["shape"], x[p], NX / Lx)
                    cy, wy = orc.weights23(s["shape"], y[p], NY / Ly)
                    for j, b in zip(cy, wy):
                        for i, a in zip(cx, wx):
                            k = ((j - 1) % NY, (i - 1) % NX)
                            if half == 0:
                                rho[k] += a * b * qw
                            else:
                                J[:, k[0], k[1]] += a * b * v * qw
        F[1], F[5], F[6], F[7] = F[0], F[2], F[3], F[4]
        F[0], F[2], F[3], F[4] = ph, ax, ay, az
        F[8] = rho
        F[9:12] = J
        # diagnose!
        pad = np.pad(EB, ((0, 0), (buffer, buffer), (buffer, buffer)), mode="wrap")
        ke = sum(np.sum(s["xyv"][2:] ** 2) * s["mass"] / 2 * s["weight"] for s in sp)
        pm = sum(np.sum(s["xyv"][2:], axis=1) * (s["mass"] * s["weight"]) for s in sp)
        pc = sum(np.sum(np.abs(s["xyv"][2:]), axis=1) * (s["mass"] * s["weight"]) for s in sp)
        e, b = F[12:15].real, F[15:18].real
        fm = np.array([np.mean(e[1] * b[2] - e[2] * b[1]), np.mean(e[2] * b[0] - e[0] * b[2]),
                       np.mean(e[0] * b[1] - e[1] * b[0])])
        rows.append(np.concatenate([[ke, np.mean(pad ** 2) / 2], pm, fm, pc]))
        for k in (2, 3, 4, 0, 8, 9, 10, 11):
            F[k] = np.fft.ifft2(F[k])
        for a, k in enumerate((12, 13, 14, 15, 16, 17, 2, 3, 4, 0, 8, 9, 10, 11)):
            z = F[k].real.reshape(NY // ngskip, ngskip, NX // ngskip, ngskip)
            bins[a, t] += z.sum(axis=(1, 3)) / ngskip ** 2
    return sp, F, np.array(rows), bins


def test_steps_against_numpy(orc):
    NX, NY, Lx, Ly, dt, buffer, ngskip, T = 16, 16, 1.0, 2.0, 0.02, 3, 2, 3
    B0 = np.array([0.6, 0.0, 0.2])
    sp = random_species(Lx, Ly, P=48, seed=8)
    sp, F = orc.lg_warmup(NX, NY, Lx, Ly, dt, sp)
    o = orc.lg_run(NX, NY, Lx, Ly, dt, B0, sp, F, T, theta=1.0, buffer=buffer, ngskip=ngskip)
    n = numpy_steps(orc, NX, NY, Lx, Ly, dt, B0, 1.0, buffer, ngskip, sp, F, T)
    for a, b in zip(o[0], n[0]):
        assert np.allclose(a["xyv"], b["xyv"], rtol=1e-10, atol=1e-13)
    scale = np.abs(n[1]).max(axis=(1, 2), keepdims=True) + 1e-300
    assert np.all(np.abs(o[1] - n[1]) <= 1e-9 * scale)
    assert np.allclose(o[2], n[2], rtol=1e-9, atol=1e-14)
    bscale = np.abs(n[3]).max(axis=(1, 2, 3), keepdims=True) + 1e-300
    assert np.all(np.abs(o[3] - n[3]) <= 1e-9 * bscale)
    # the dynamics is not trivial: fields and currents are non-zero
    assert np.abs(o[1][12:]).max() > 0 and np.abs(o[1][9:12]).max() > 0


def test_two_calls_equal_one(orc):
    NX, NY, Lx, Ly, dt = 16, 8, 1.0, 1.0, 0.02
    B0 = np.array([0.6, 0.0, 0.0])
    sp = random_species(Lx, Ly, P=32, seed=9)
    sp, F = orc.lg_warmup(NX, NY, Lx, Ly, dt, sp)
    whole = orc.lg_run(NX, NY, Lx, Ly, dt, B0, sp, F, 6, ntskip=2, ngskip=2, buffer=2)
    a = orc.lg_run(NX, NY, Lx, Ly, dt, B0, sp, F, 3, ntskip=2, ngskip=2, buffer=2, nbins=3)
    b = orc.lg_run(NX, NY, Lx, Ly, dt, B0, a[0], a[1], 3, t0=3, ntskip=2, ngskip=2, buffer=2, nbins=3, bins=a[3])
    for p, q in zip(b[0], whole[0]):
        assert np.array_equal(p["xyv"], q["xyv"])
    assert np.array_equal(b[1], whole[1])
    assert np.array_equal(np.concatenate([a[2], b[2]]), whole[2])
    assert np.array_equal(b[3], whole[3])


# ---- LorenzGaugeField (the centred scheme, :665-745) ------------------------

def test_centred_single_mode(orc):
    """phi+ = cos, phi0 = 0, no source: the new phi+ is r cos and phi0 the old
    phi+, so Ex = -d((phi+ + phi0)/2)/dx = (r + 1)/2 k sin; Az+ = cos gives
    Ez = -(r - 1)/dt cos and By = (r + 1)/2 k sin."""
    NX, NY, Lx, Ly, dt = 16, 8, 2.0, 0.5, 0.01
    for theta in (0.0, 1.0):
        for mx, my in ((1, 0), (3, 0), (0, 2)):
            kx0, ky0 = 2 * np.pi * mx / Lx, 2 * np.pi * my / Ly
            k2 = kx0 ** 2 + ky0 ** 2
            r = (2 - dt ** 2 * k2 * (1 - theta)) / (1 + dt ** 2 * k2 * theta / 2)
            c, s = mode(NX, NY, Lx, Ly, mx, my), mode(NX, NY, Lx, Ly, mx, my, "sin")
            F = np.zeros((18, NY, NX), dtype=complex)
            F[0] = np.fft.fft2(c)
            F1 = orc.lg_solve_centred(F, dt, theta, Lx, Ly)
            assert np.allclose(F1[0], r * F[0], atol=1e-11) and np.array_equal(F1[1], F[0])
            assert np.allclose(F1[12].real, (r + 1) / 2 * kx0 * s, atol=1e-11)
            assert np.allclose(F1[13].real, (r + 1) / 2 * ky0 * s, atol=1e-11)
            F = np.zeros((18, NY, NX), dtype=complex)
            F[4] = np.fft.fft2(c)
            F1 = orc.lg_solve_centred(F, dt, theta, Lx, Ly)
            assert np.allclose(F1[14].real, -(r - 1) / dt * c, atol=1e-9)
            assert np.allclose(F1[16].real, (r + 1) / 2 * kx0 * s, atol=1e-11)


@pytest.mark.parametrize("theta,dt", [(1.0, 0.05), (0.0, 0.005)])
def test_centred_vacuum_recurrence(orc, theta, dt):
    NX, NY, Lx, Ly, T, m = 16, 8, 1.0, 1.0, 50, 5
    k2 = (2 * np.pi * m / Lx) ** 2
    r = (2 - dt ** 2 * k2 * (1 - theta)) / (1 + dt ** 2 * k2 * theta / 2)
    F = np.zeros((18, NY, NX), dtype=complex)
    F[0, 0, m] = 1.0
    F[1, 0, m] = 0.5
    sp = inert_species(Lx, Ly)
    _, F1, _ = orc.lg_run_centred(NX, NY, Lx, Ly, dt, [0, 0, 0], sp, F, np.zeros((6, NY, NX)), T, theta=theta)
    s = np.linalg.matrix_power(np.array([[r, -1.0], [1.0, 0.0]]), T) @ np.array([1.0, 0.5])
    assert np.allclose([F1[0, 0, m], F1[1, 0, m]], s, rtol=1e-9, atol=1e-12)
    # the sources were transformed and their [1, 1] cleared
    assert not np.any(F1[8:12])


def test_centred_particles_move_before_the_first_update(orc):
    """The particle loop comes first and EBxyz is zero in a fresh field:
    uncharged particles stream in the first step and turn about B0 only in
    the second (the Boris turn 2 atan(|t|))."""
    NX, NY, Lx, Ly, dt = 16, 16, 1.0, 1.0, 0.01
    B0 = np.array([0.0, 0.0, 3.0])
    sp = inert_species(Lx, Ly, P=8, seed=11)
    x0 = sp[0]["xyv"].copy()
    F = np.zeros((18, NY, NX), dtype=complex)
    sp1, F1, EB1 = orc.lg_run_centred(NX, NY, Lx, Ly, dt, B0, sp, F, np.zeros((6, NY, NX)), 1)
    v = x0[2:]
    assert np.array_equal(sp1[0]["xyv"][2:], v)
    assert np.array_equal(sp1[0]["xyv"][0], [orc.unimod23(a + (b + b) / 2 * dt, Lx) for a, b in zip(x0[0], v[0])])
    assert np.array_equal(EB1[5], np.full((NY, NX), 3.0)) and not np.any(EB1[:5])
    sp2, _, _ = orc.lg_run_centred(NX, NY, Lx, Ly, dt, B0, sp1, F1, EB1, 1)
    S = np.array([orc.weights23("bspline2", x, NX / Lx)[1] for x in sp1[0]["xyv"][0]]).sum(axis=1)
    S *= np.array([sum(orc.weights23("bspline2", y, NY / Ly)[1]) for y in sp1[0]["xyv"][1]])
    phi = 2 * np.arctan(3.0 * S * dt / 2)
    z0 = v[0] + 1j * v[1]
    assert np.allclose(sp2[0]["xyv"][2] + 1j * sp2[0]["xyv"][3], z0 * np.exp(-1j * phi), rtol=0, atol=1e-14)


def numpy_steps_centred(orc, NX, NY, Lx, Ly, dt, B0, theta, species, F, EB, T):
    kx = (2 * np.pi / Lx) * np.concatenate([np.arange(NX // 2), np.arange(-NX // 2, 0)])
    ky = (2 * np.pi / Ly) * np.concatenate([np.arange(NY // 2), np.arange(-NY // 2, 0)])
    KX, KY = kx[None, :], ky[:, None]
    k2 = KX ** 2 + KY ** 2
    dV = (Lx / NX) * (Ly / NY)
    F, EB = F.copy(), EB.copy()
    sp = [dict(s, xyv=s["xyv"].copy()) for s in species]
    for _ in range(T):
        rho, J = np.zeros((NY, NX)), np.zeros((3, NY, NX))
        for s in sp:
            x, y, vx, vy, vz = s["xyv"]
            qw = s["charge"] * s["weight"] / dV
            q_m = s["charge"] / s["mass"]
            for p in range(x.size):
                g = np.array([orc.gather23(EB[c], EB[c + 1], x[p], y[p], s["shape"], Lx, Ly) for c in (0, 2, 4)]).ravel()
                u = np.array([vx[p], vy[p], vz[p]])
                v = numpy_boris(dt, q_m, g[:3], g[3:], u)
                vx[p], vy[p], vz[p] = v
                x[p] = orc.unimod23(x[p] + (u[0] + v[0]) / 2 * dt, Lx)
                y[p] = orc.unimod23(y[p] + (u[1] + v[1]) / 2 * dt, Ly)
                cx, wx = orc.weights23(s["shape"], x[p], NX / Lx)
                cy, wy = orc.weights23(s["shape"], y[p], NY / Ly)
                for j, b in zip(cy, wy):
                    for i, a in zip(cx, wx):
                        k = ((j - 1) % NY, (i - 1) % NX)
                        rho[k] += a * b * qw
                        J[:, k[0], k[1]] += a * b * v * qw
        src = [np.fft.fft2(rho)] + [np.fft.fft2(J[c]) for c in range(3)]
        for a in src:
            a[0, 0] = 0
        num = 2 - dt ** 2 * k2 * (1 - theta)
        den = 1 + dt ** 2 * k2 * theta / 2
        for a, (xp, x0) in enumerate(((0, 1), (2, 5), (3, 6), (4, 7))):
            new = (num * F[xp] + dt ** 2 * src[a]) / den - F[x0]
            F[x0] = F[xp]
            F[xp] = new
        F[8:12] = src
        ph = F[0] + F[1]
        E = [-1j * KX * ph / 2 - (F[2] - F[5]) / dt, -1j * KY * ph / 2 - (F[3] - F[6]) / dt, -(F[4] - F[7]) / dt]
        B = [1j * KY * (F[4] + F[7]) / 2, -1j * KX * (F[4] + F[7]) / 2,
             1j * KX * (F[3] + F[6]) / 2 - 1j * KY * (F[2] + F[5]) / 2]
        for c in range(3):
            F[12 + c] = np.fft.ifft2(E[c])
            F[15 + c] = np.fft.ifft2(B[c])
        EB = np.array([F[12 + c].real + (B0[c - 3] if c >= 3 else 0.0) for c in range(6)])
    return sp, F, EB


def test_centred_steps_against_numpy(orc):
    NX, NY, Lx, Ly, dt, T = 16, 16, 1.0, 2.0, 0.02, 3
    B0 = np.array([0.6, 0.0, 0.2])
    sp = random_species(Lx, Ly, P=48, seed=12)
    F, EB = np.zeros((18, NY, NX), dtype=complex), np.zeros((6, NY, NX))
    o = orc.lg_run_centred(NX, NY, Lx, Ly, dt, B0, sp, F, EB, T, theta=1.0)
    n = numpy_steps_centred(orc, NX, NY, Lx, Ly, dt, B0, 1.0, sp, F, EB, T)
    for a, b in zip(o[0], n[0]):
        assert np.allclose(a["xyv"], b["xyv"], rtol=1e-10, atol=1e-13)
    scale = np.abs(n[1]).max(axis=(1, 2), keepdims=True) + 1e-300
    assert np.all(np.abs(o[1] - n[1]) <= 1e-9 * scale)
    assert np.allclose(o[2], n[2], rtol=1e-9, atol=1e-12)
    assert np.abs(o[1][12:]).max() > 0
    # two calls equal one
    a = orc.lg_run_centred(NX, NY, Lx, Ly, dt, B0, sp, F, EB, 1, theta=1.0)
    b = orc.lg_run_centred(NX, NY, Lx, Ly, dt, B0, a[0], a[1], a[2], T - 1, theta=1.0)
    assert np.array_equal(b[1], o[1]) and np.array_equal(b[2], o[2])
    for p, q in zip(b[0], o[0]):
        assert np.array_equal(p["xyv"], q["xyv"])
