"""Pins the magnetised 1D2V restatement (oracle/pic1d2v.c) of src/NGP1D2V.jl
and src/NGP1D2V2S.jl: the Boris rotation against closed forms and the
textbook vector form, the signed exact deposit against exact rationals, a
transcription of the scripts' loop over the pinned primitives, and the plasma
and upper-hybrid frequencies of a cold (magnetised) plasma.""" 
from fractions import Fraction

import numpy as np
import pytest


def test_boris12_rotation_with_no_field(orc):
    rng = np.random.default_rng(0)
    for _ in range(200):
        vx, vy = rng.standard_normal(2)
        B, dt, qm = rng.uniform(0.1, 5.0), rng.uniform(1e-3, 0.5), rng.choice([1.0, -1.0, 0.125])
        ux, uy = orc.boris12(vx, vy, 0.0, B, dt, qm)
        tau = B * dt / 2 * qm
        th = 2 * np.arctan(tau)                       # clockwise for q_m B > 0
        ex, ey = vx * np.cos(th) + vy * np.sin(th), -vx * np.sin(th) + vy * np.cos(th)
        assert ux == pytest.approx(ex, abs=1e-14) and uy == pytest.approx(ey, abs=1e-14)
        assert np.hypot(ux, uy) == pytest.approx(np.hypot(vx, vy), rel=1e-14)


def test_boris12_against_the_vector_form(orc):
    """v- = v + qm E dt/2 x, v+ = v- + 2 (v- + v- x t) x t / (1 + t.t), v = v+ + qm E dt/2 x,
    evaluated in long double."""
    L = np.longdouble
    rng = np.random.default_rng(1)
    for _ in range(200):
        vx, vy, E = rng.standard_normal(3)
        B, dt, qm = rng.uniform(0.1, 5.0), rng.uniform(1e-3, 0.5), rng.choice([1.0, -1.0, 0.125])
        h = L(dt) / 2 * L(qm)
        vm = np.array([L(vx) + L(E) * h, L(vy), L(0)])
        t = np.array([L(0), L(0), L(B) * h])
        a = vm + np.cross(vm, t)
        vp = vm + 2 * np.cross(a, t) / (1 + t.dot(t))
        ux, uy = orc.boris12(vx, vy, E, B, dt, qm)
        assert ux == pytest.approx(float(vp[0] + L(E) * h), rel=1e-14, abs=1e-15)
        assert uy == pytest.approx(float(vp[1]), rel=1e-14, abs=1e-15)


def test_signed_deposit_is_the_exact_sum(orc):
    N, R, P = 16, 7, 40
    rng = np.random.default_rng(2)
    xs = [rng.random(P), rng.random(P)]
    q, w = (-1.0, 1.0), 0.37
    got = orc.deposit_1d2v(N, R, q, w, xs)
    exact = [Fraction(0)] * N
    for s in range(2):
        for c in xs[s]:
            cells, wt = orc.gauss_stencil(N, R, c)
            for k in range(2 * R + 1):
                exact[cells[k]] += Fraction((q[s] * wt[k]) * w)
    assert np.array_equal(got, np.array([float(e) for e in exact]))
    # one species of charge +1 is the scripts' plain deposit
    assert np.array_equal(orc.deposit_1d2v(N, R, (1.0,), w, xs[:1]), orc.deposit(N, R, xs[0], w=w))


def transcribed(orc, N, R, q, qm, mass, w, n0, B0, dt, bin, xs, vxs, vys, T):
    """src/NGP1D2V2S.jl:31-58 line by line over the pinned primitives."""
    ns, P = len(xs), xs[0].size
    xs, vxs, vys = [a.copy() for a in xs], [a.copy() for a in vxs], [a.copy() for a in vys]
    nb = (T + bin - 1) // bin
    Es, D = np.zeros((nb, N)), np.zeros((nb, 5))
    for t in range(1, T + 1):
        E = orc.fieldsolve(orc.deposit_1d2v(N, R, q, w, xs))
        for s in range(ns):
            for j in range(P):
                Ej = orc.gather_at(N, R, E, xs[s][j])
                vxs[s][j], vys[s][j] = orc.boris12(vxs[s][j], vys[s][j], Ej, B0, dt, qm[s])
                xs[s][j] += vxs[s][j] * dt
        xs = [orc.wrap(a) for a in xs]
        b = -(-t // bin) - 1
        Es[b] += E
        if t % bin == 0:
            kin = mass[0] * (vys[0] * vys[0] + vxs[0] * vxs[0])
            if ns == 2:
                kin = kin + mass[1] * (vys[1] * vys[1] + vxs[1] * vxs[1])
            D[b, 0] += (orc.tree_sum(E * E) / N) / 2
            D[b, 1] += ((orc.tree_sum(kin) * n0) / P) / 2
            for c, vv in ((3, vxs), (4, vys)):
                m = (mass[0] * vv[0]) / P
                if ns == 2:
                    m = m + (mass[1] * vv[1]) / P
                D[b, c] += orc.tree_sum(m)
            D[b, 2] += D[b, 0] + D[b, 1]
            D[b, :3] *= 2 / n0
    return xs, vxs, vys, Es, D


@pytest.mark.parametrize("ns", [1, 2])
def test_run_matches_the_transcribed_loop(orc, ns):
    N, R, P, T, bin = 32, 7, 200, 9, 3
    n0 = 4 * np.pi ** 2
    vth = np.sqrt(n0) / N / 4
    B0 = np.sqrt(n0) / 16
    dt = 1 / N / 6 / vth
    rng = np.random.default_rng(ns)
    if ns == 1:
        q, qm, mass, w = (1.0,), (1.0,), (1.0,), n0 / P
    else:
        q, qm, mass, w = (-1.0, 1.0), (-1.0, 1 / 8), (1.0, 8.0), n0 / (2 * P)
    xs = [rng.random(P) for _ in range(ns)]
    vxs = [vth * rng.standard_normal(P) for _ in range(ns)]
    vys = [vth * rng.standard_normal(P) for _ in range(ns)]
    ref = transcribed(orc, N, R, q, qm, mass, w, n0, B0, dt, bin, xs, vxs, vys, T)
    for nthreads in (1, 4):
        got = orc.run_1d2v(N, R, q, qm, mass, w, n0, B0, dt, bin, xs, vxs, vys, T, nthreads=nthreads)
        for g, r in zip(got, ref):
            if isinstance(g, list):
                assert all(np.array_equal(a, b) for a, b in zip(g, r))
            else:
                assert np.array_equal(g, r)
    # the same run in two calls (t0 carries the step count and the bins)
    a = orc.run_1d2v(N, R, q, qm, mass, w, n0, B0, dt, bin, xs, vxs, vys, 4, nbins=3)
    b = orc.run_1d2v(N, R, q, qm, mass, w, n0, B0, dt, bin, a[0], a[1], a[2], T - 4, t0=4, Es=a[3], D=a[4])
    for g, r in zip(b, ref):
        if isinstance(g, list):
            assert all(np.array_equal(x, y) for x, y in zip(g, r))
        else:
            assert np.array_equal(g, r)


@pytest.mark.parametrize("B0", [0.0, 2 * np.pi])
def test_cold_magnetised_plasma_oscillates_at_the_upper_hybrid_frequency(orc, B0):
    # one cold species on a perturbed lattice: omega^2 = wp^2 S(k)^2 + B0^2,
    # wp^2 = P w / N (the per-cell charge of the uniform plasma; the scripts'
    # w = n0/P makes that n0/N)
    N, R = 64, 7
    P = 32 * N
    n0 = 4 * np.pi ** 2
    w = n0 / P * N                                    # wp = sqrt(n0) = 2 pi
    dt = 1 / (6 * N)
    x0 = (np.arange(P) + 0.5) / P
    x = x0 + 1e-4 * np.sin(2 * np.pi * x0)
    z = np.zeros(P)
    T = 1200
    _, _, _, Es, D = orc.run_1d2v(N, R, (1.0,), (1.0,), (1.0,), w, n0, B0, dt, 1, [x], [z], [z], T, nthreads=8)
    d1 = D[:, 0]
    mins = [i for i in range(1, T - 1) if d1[i] < d1[i - 1] and d1[i] <= d1[i + 1]]
    # the displacement (and E) oscillates about x0 B0^2/omega^2, started at rest:
    # it crosses zero, and E^2 dips, twice a period only without the field
    omega = (np.pi if B0 == 0 else 2 * np.pi) / (np.mean(np.diff(mins)) * dt)
    k = 2 * np.pi
    S = np.exp(-k ** 2 / (4 * N ** 2)) * np.sin(k / (2 * N)) / (k / (2 * N))
    assert omega == pytest.approx(np.sqrt(n0 * S ** 2 + B0 ** 2), rel=3e-3)
    # one bin per step: Es holds each step's field, D1 its energy
    assert Es.shape == (T, N)
    assert np.allclose(np.sum(Es ** 2, axis=1) / N / 2 * (2 / n0), d1, rtol=1e-12)
