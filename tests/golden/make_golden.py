"""Writes the full-run golden fixtures of the two small workloads from the CPU
restatement (the oracle, oracle/picoracle.c), in the build container:

  quiet_two_stream_8192.npz  N=64, P=2048, Gaussian +-7 midpoint, dt=1/6N,
                             rtol=4eps, bit-reversal quiet start, 8192 steps
                             (src/GaussianFixedPointQuiet.jl:1-13)
  ngp_two_stream_1024.npz    N=128, P=8192, NGP leapfrog, dt=1/4N, W=200,
                             seeded uniform x, v=-1/+1 by half, 1024 steps
                             (src/NGPFourier.jl:1-6)

Each holds the sweep counts of every step, the D[t,1:4] trace every `stride`
steps, the final E and x, v (every `pstride`-th particle), and SHA-256 digests of the whole D trace and of
the final state, so a GPU run is compared with the oracle bit for bit over the
full run without the oracle running on the GPU host.

    python tests/golden/make_golden.py
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


def main():
    from picgolf import configs
    cfg, x, v, T = configs.quiet_two_stream()
    write("quiet_two_stream_8192.npz", cfg, 7, 1, x, v, T, 8)
    cfg, x, v, T = configs.ngp_two_stream()
    write("ngp_two_stream_1024.npz", cfg, 0, 0, x, v, T, 4, pstride=16)


if __name__ == "__main__":
    main()
