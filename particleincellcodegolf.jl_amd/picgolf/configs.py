"""The five benchmark workloads with their initial conditions.

Initial particle states are generated once on the host (numpy PCG64 streams
from a seed) so the GPU run and the CPU restatement start from identical bits.

  ngp_two_stream     N=128, P=8192, NGP leapfrog, dt=1/4N, 1024 steps, W=200
                     (src/NGPFourier.jl:1-2; seeded uniform x, v=-1/+1 by half)
  quiet_two_stream   N=64, P=2048, Gaussian +-7 midpoint, dt=1/6N, 8192 steps,
                     W=32pi^2/3, rtol=4eps, bit-reversal quiet start
                     (src/GaussianFixedPointQuiet.jl:1-6)
  quiet_simpson      the same with the Simpson-1/3 fixed point, Gaussian or
                     area shape (src/*FixedPointQuietSimpson13.jl:1-5)
  scaled_two_stream  N=4096, P=2^24, Gaussian +-7 midpoint, dt=1/6N, 2000 steps,
                     W=400, rtol=1e-8 (src/GaussianFixedPoint.jl:4), seeded
                     uniform x, v=-1/+1 by half plus N(0, 0.05^2)
  ensemble           512 instances: N in {32..1024}, 16-256 particles per cell,
                     W in [100, 400], beams +-u with u in [0.5, 1.5]
  bernstein          magnetised 1D2V, one species (N=512, P=15N) or two
                     (N=256, P=8N each, M=8), Boris push with the Gaussian
                     +-7 stencil (src/NGP1D2V.jl, src/NGP1D2V2S.jl)
  es2d3v             256 x 256 grid, 64 particles per cell, electrostatic 2D3V
                     with B = B0 x (src/Electrostatic2D3V.jl): seeded uniform
                     positions in (0, 1], vx = -1/+1 by half + N(0, 0.1^2),
                     vy, vz ~ N(0, 0.1^2); dt = 1/1024, 500 steps, n0 = 4 pi^2,
                     B0 = sqrt(n0)/4, w = n0/P/(dx dy)
"""
import numpy as np

from .core import Config
from .core2d import Config2D

EPS = float(np.finfo(np.float64).eps)


def quiet_start(P):
    """x = (bitreverse.(0:P-1) .+ 2.0^63) / 2.0^64 over signed Int64;
    v = -1 for the first half, +1 after (src/GaussianFixedPointQuiet.jl:2-3)."""
    i = np.arange(P, dtype=np.uint64)
    r = np.zeros(P, dtype=np.uint64)
    for b in range(64):
        r |= ((i >> np.uint64(b)) & np.uint64(1)) << np.uint64(63 - b)
    x = (r.view(np.int64).astype(np.float64) + 2.0 ** 63) / 2.0 ** 64
    v = np.where(np.arange(1, P + 1) > P / 2, 1.0, -1.0)
    return x, v


def beams(P, u=1.0):
    """v = collect(1:P .> P/2) .* 2 .- 1 scaled by the beam speed u."""
    return np.where(np.arange(1, P + 1) > P / 2, u, -u).astype(np.float64)


def ngp_two_stream(seed=0, P=None):
    N = 128
    P = 64 * N if P is None else P
    W = 200.0
    cfg = Config(N=N, P=P, W=W, w=W / P * N, dt=1 / (4 * N), shape="ngp", scheme="leapfrog")
    rng = np.random.default_rng(seed)
    return cfg, rng.random(P), beams(P), 1024


def quiet_two_stream(precision="fp64"):
    N = 64
    P = 32 * N
    W = 32 * np.pi ** 2 / 3
    cfg = Config(N=N, P=P, W=W, w=W / P * N, dt=1 / (6 * N), shape="gauss7", scheme="midpoint",
                 precision=precision, rtol=4 * EPS, max_sweeps=10)
    x, v = quiet_start(P)
    return cfg, x, v, 2 ** 13


def quiet_simpson(shape="gauss7", precision="fp64"):
    """The Simpson-1/3 fixed-point scripts: N=64, P=2048, dt=1/6N, 8192 steps,
    W=32pi^2/3, bit-reversal quiet start; the Gaussian +-7 shape with
    rtol=4eps (src/GaussianFixedPointQuietSimpson13.jl:1-5) or the linear area
    weights with rtol=1e-14 (src/AreaFixedPointQuietSimpson13.jl:1-5)."""
    N = 64
    P = 32 * N
    W = 32 * np.pi ** 2 / 3
    rtol = 1e-14 if shape == "area" else 4 * EPS
    cfg = Config(N=N, P=P, W=W, w=W / P * N, dt=1 / (6 * N), shape=shape, scheme="simpson",
                 precision=precision, rtol=rtol, max_sweeps=10)
    x, v = quiet_start(P)
    return cfg, x, v, 2 ** 13


def bernstein(species=1, seed=0):
    """The magnetised 1D2V scripts (Bernstein waves).  species=1:
    src/NGP1D2V.jl:22-32 -- N=512, P=15N, T=2^14 in bins of T/TO = 16 steps,
    n0 = 4pi^2, vth = sqrt(n0)/N/4, dt = 1/N/6vth (6vth binds first), B0 = sqrt(n0)/16,
    w = n0/P.  species=2: src/NGP1D2V2S.jl:13-21 -- N=256, P=8N per species,
    T=2^16 in bins of 32 steps, M=8, vth = sqrt(n0)/N/8, dt = 1/N/16vth,
    B0 = sqrt(n0)/8, w = n0/2P, charges -1 and +1, q/m -1 and 1/M; positions
    from the bit-reversal quiet start (the second species shifted by pi).
    Velocities are vth*erfinv(U) as in the scripts (U uniform in [0, 1), so
    every component is >= 0); the uniforms come from a seeded numpy stream
    (species=1) or the scripts' additive recurrences (species=2)."""
    from scipy.special import erfinv
    from .core1d2v import Config1D2V
    n0 = 4 * np.pi ** 2
    if species == 1:
        N = 512
        P = 15 * N
        T, TO = 2 ** 14, 2 ** 14 // 16
        vth = np.sqrt(n0) / N / 4
        cfg = Config1D2V(N=N, P=P, bin=T // TO, w=n0 / P, n0=n0, B0=np.sqrt(n0) / 16, dt=1 / N / (6 * vth))
        rng = np.random.default_rng(seed)
        x, rvx, rvy = rng.random(P), rng.random(P), rng.random(P)
        return cfg, [(x, vth * erfinv(rvx), vth * erfinv(rvy))], T
    N = 256
    P = 8 * N
    T, TO, M = 2 ** 16, 2 ** 16 // 32, 8
    vth = np.sqrt(n0) / N / 8
    cfg = Config1D2V(N=N, P=P, bin=T // TO, w=n0 / (2 * P), n0=n0, B0=np.sqrt(n0) / 8, dt=1 / N / (16 * vth),
                     q=(-1.0, 1.0), q_m=(-1.0, 1 / M), mass=(1.0, float(M)))
    xq, _ = quiet_start(P)
    x1 = xq - np.floor(xq)
    x2 = (xq + np.pi) - np.floor(xq + np.pi)
    # R(alpha, P) = rem.(n*alpha, 1), n = 1..P, alpha = 1/root of x^(d+1) = x + 1
    def golden(d):
        r = 1.5
        for _ in range(200):
            r = r - (r ** (d + 1) - r - 1) / ((d + 1) * r ** d - 1)
        return 1 / r
    n = np.arange(1, P + 1)
    vx1 = vth * erfinv(np.remainder(n * golden(1), 1.0))
    vy1 = vth * erfinv(np.remainder(n * golden(2), 1.0))
    return cfg, [(x1, vx1, vy1), (x2, vx1 / np.sqrt(M), vy1 / np.sqrt(M))], T


def scaled_two_stream(precision="mixed", seed=0, P=2 ** 24, N=4096):
    W = 400.0
    cfg = Config(N=N, P=P, W=W, w=W / P * N, dt=1 / (6 * N), shape="gauss7", scheme="midpoint",
                 precision=precision, rtol=1e-8, max_sweeps=10)
    rng = np.random.default_rng(seed)
    x = rng.random(P)
    v = beams(P) + 0.05 * rng.standard_normal(P)
    return cfg, x, v, 2000


def ensemble(n=512, precision="mixed", seed=0):
    """Independent two-stream instances with uneven N, P, W and beam speed."""
    rng = np.random.default_rng(seed)
    cfgs, states = [], []
    for k in range(n):
        N = int(2 ** rng.integers(5, 11))        # 32 .. 1024
        ppc = int(rng.integers(16, 257))
        P = N * ppc
        W = float(rng.uniform(100.0, 400.0))
        u = float(rng.uniform(0.5, 1.5))
        cfgs.append(Config(N=N, P=P, W=W, w=W / P * N, dt=1 / (6 * N), shape="gauss7", scheme="midpoint",
                           precision=precision, rtol=1e-8, max_sweeps=10))
        r = np.random.default_rng(seed * 100003 + k + 1)
        states.append((r.random(P), beams(P, u)))
    return cfgs, states, 1000


def es2d3v(precision="mixed", seed=0, NX=256, NY=256, ppc=64, pairing="reference"):
    """The electrostatic 2D3V workload.  Positions are drawn in (0, 1] (the
    script's unimod convention); n0 = 4 pi^2 and B0 = sqrt(n0)/4 as in
    src/Electrostatic2D3V.jl:24."""
    P = NX * NY * ppc
    n0 = 4 * np.pi ** 2
    cfg = Config2D(NX=NX, NY=NY, P=P, dt=1 / 1024, w=n0 / P * NX * NY, B0=np.sqrt(n0) / 4, ns=2,
                   precision=precision, pairing=pairing)
    rng = np.random.default_rng(seed)
    x = 1.0 - rng.random(P)          # (0, 1]
    y = 1.0 - rng.random(P)
    vx = beams(P) + 0.1 * rng.standard_normal(P)
    vy = 0.1 * rng.standard_normal(P)
    vz = 0.1 * rng.standard_normal(P)
    return cfg, (x, y, vx, vy, vz), 500
