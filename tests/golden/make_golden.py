"""Writes the full-run golden fixtures of the two small workloads from the CPU
restatement (the oracle, oracle/picoracle.c), in the build container:

  quiet_two_stream_8192.npz  N=64, P=2048, Gaussian +-7 midpoint, dt=1/6N,
                             rtol=4eps, bit-reversal quiet start, 8192 steps
                             (src/GaussianFixedPointQuiet.jl:1-13)
  ngp_two_stream_1024.npz    N=128, P=8192, NGP leapfrog, dt=1/4N, W=200,
                             seeded uniform x, v=-1/+1 by half, 1024 steps
                             (src/NGPFourier.jl:1-6)
  scaled_two_stream_p65536_2000.npz
                             the scaled two-stream workload (N=4096, Gaussian
                             +-7 midpoint, rtol=1e-8, dt=1/6N, W=400, seeded
                             x, +-1 beams + N(0, 0.05^2)) at P=2^16 over its
                             full 2000 steps (src/GaussianFixedPoint.jl)
  es2d3v_256_500.npz         the electrostatic 2D3V workload, 256 x 256 cells,
                             64 particles per cell, the script's shape pairing,
                             FP64, 500 steps (src/Electrostatic2D3V.jl:24-57);
                             about 5 minutes on 8 threads

Each holds the sweep counts of every step, the D[t,1:4] trace every `stride`
steps, the final E and x, v (every `pstride`-th particle), and SHA-256 digests of the whole D trace and of
the final state, so a GPU run is compared with the oracle bit for bit over the
full run without the oracle running on the GPU host.

    python tests/golden/make_golden.py [1d] [scaled] [2d]
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "particleincellcodegolf.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def write(name, cfg, shape, scheme, x, v, nsteps, stride, pstride=1):
    import oracle
    xo, vo, Eo, Do, So = oracle.run(cfg.N, shape, scheme, cfg.W, cfg.w, cfg.dt, x, v, nsteps, rtol=cfg.rtol,
                                    max_sweeps=cfg.max_sweeps, nthreads=8)
    path = os.path.join(HERE, name)
    np.savez_compressed(path, sweeps=So.astype(np.int8), D_every=Do[stride - 1::stride], stride=stride,
                        nsteps=nsteps, pstride=pstride, x=xo[::pstride], v=vo[::pstride], E=Eo, D_sha256=digest(Do), S_sha256=digest(So.astype(np.int32)),
                        x_sha256=digest(xo), v_sha256=digest(vo), E_sha256=digest(Eo))
    print(f"{path}: {os.path.getsize(path)} bytes, mean sweeps {So.mean():.3f}")


def write2d(name, cfg, st, nsteps, pstride):
    import oracle
    x, y, vx, vy, vz, Ex, Ey, K = oracle.run2d(cfg.NX, cfg.NY, cfg.dt, cfg.B0, cfg.w, *st, nsteps, ns=cfg.ns,
                                               textbook=1 if cfg.pairing == "textbook" else 0, nthreads=8)
    parts = {"x": x, "y": y, "vx": vx, "vy": vy, "vz": vz}
    path = os.path.join(HERE, name)
    np.savez_compressed(path, nsteps=nsteps, pstride=pstride, K=K, K_sha256=digest(K),
                        Ex_sha256=digest(Ex), Ey_sha256=digest(Ey), Ex_row=Ex[::37].copy(),
                        **{k: a[::pstride] for k, a in parts.items()},
                        **{k + "_sha256": digest(a) for k, a in parts.items()})
    print(f"{path}: {os.path.getsize(path)} bytes")


def main(which=("1d", "scaled", "2d")):
    from picgolf import configs
    if "2d" in which:
        cfg, st, T = configs.es2d3v(precision="fp64")
        write2d("es2d3v_256_500.npz", cfg, st, T, pstride=4099)
    if "scaled" in which:
        cfg, x, v, T = configs.scaled_two_stream(precision="fp64", P=2 ** 16)
        write("scaled_two_stream_p65536_2000.npz", cfg, 7, 1, x, v, T, 20, pstride=64)
    if "1d" not in which:
        return
    cfg, x, v, T = configs.quiet_two_stream()
    write("quiet_two_stream_8192.npz", cfg, 7, 1, x, v, T, 8)
    cfg, x, v, T = configs.ngp_two_stream()
    write("ngp_two_stream_1024.npz", cfg, 0, 0, x, v, T, 4, pstride=16)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("1d", "scaled", "2d"))
