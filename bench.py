"""Benchmark: particle-pushes/s of the fixed-point midpoint PIC step on one MI355X.

Workload: the scaled two-stream run -- N=4096 cells, P=2^24 particles,
erf-integrated Gaussian +-7 stencil, implicit midpoint with <=10 Picard
sweeps at rtol=1e-8, dt=1/6N, W=400 -- in mixed precision (the headline
`value`) and in plain FP64 (reported beside it).  1 push = 1 particle advanced
one full time step including all sweeps.  Inputs are resident in HBM before
the timed region; K steps are timed between two device synchronisations.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision both|mixed|fp64]
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "particleincellcodegolf.jl_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--particles", type=int, default=2 ** 24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the 2000-step mixed-vs-FP64 accuracy run of the headline workload")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the ensemble and 2D3V workloads reported beside the headline")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--workload", default="scaled",
                    choices=["scaled", "small", "ensemble", "es2d3v", "pic2d3v", "lorenz2d3v", "lorenz2d3v_si", "em1d",
                             "em1d_explicit"],
                    help="scaled: the headline line (default); ensemble / es2d3v: only that workload, "
                         "--steps timed steps, one JSON line (for profiling)")
    ap.add_argument("--script-steps", type=int, default=4048,
                    help="em1d, em1d_explicit: steps of the script-sized run (the scripts' T = 4048)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default particle tiling instead of autotuning it before the timed steps")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (HBM counter calibration)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print a per-kernel table from a rocprofv3 --kernel-trace csv directory")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="per-kernel HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                         "csv directories, scaled by the same counters of --copy-calibration")
    return ap.parse_args()


# Dispatches shorter than this did no particle work: the sweep loop runs on the
# device, and launches past the converged sweep return at their first test.
PREDICATED_OFF_US = 8.0
COPY_BYTES = 1 << 30


def _short(name):
    import re
    return re.sub(r"\(.*", "", name).replace("void ", "")


def _csv_rows(d, suffix):
    import csv
    import glob
    rows = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    if not rows:
        sys.exit(f"no *{suffix} under {d}")
    return rows


def summarize_trace(d):
    """Per-kernel durations from a kernel trace, split into dispatches that did
    work and the predicated-off ones."""
    by = {}
    for r in _csv_rows(d, "kernel_trace.csv"):
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        by.setdefault(_short(r["Kernel_Name"]), []).append(us)
    print(f"dispatches >= {PREDICATED_OFF_US} us are 'live' (did particle or field work); shorter ones are "
          f"predicated off (or tiny by nature, e.g. k_diag_final)")
    print(f"{'kernel':34s} {'calls':>6s} {'avg_us':>9s} | {'live':>5s} {'live_avg_us':>11s} {'min':>8s} {'max':>8s}"
          f" | {'off':>4s} {'off_avg_us':>10s}")
    for k, ds in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        live = [x for x in ds if x >= PREDICATED_OFF_US]
        off = [x for x in ds if x < PREDICATED_OFF_US]
        la = sum(live) / len(live) if live else 0.0
        oa = sum(off) / len(off) if off else 0.0
        print(f"{k[:34]:34s} {len(ds):6d} {sum(ds) / len(ds):9.2f} | {len(live):5d} {la:11.2f} "
              f"{min(live or [0]):8.2f} {max(live or [0]):8.2f} | {len(off):4d} {oa:10.2f}")


def _pmc(d, counter):
    per, name = {}, {}
    for r in _csv_rows(d, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        key = r["Dispatch_Id"]
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
        name[key] = _short(r["Kernel_Name"])
    by = {}
    for key, v in per.items():
        by.setdefault(name[key], []).append(v)
    return by


def summarize_pmc(fetch_dir, write_dir, copy_fetch, copy_write):
    """HBM bytes per dispatch: FETCH_SIZE and WRITE_SIZE (KiB) scaled so that
    the copy kernel of known size reads and writes exactly COPY_BYTES."""
    cfs = [v for k, l in _pmc(copy_fetch, "FETCH_SIZE").items() for v in l]
    cws = [v for k, l in _pmc(copy_write, "WRITE_SIZE").items() for v in l]
    # the copy is the dispatch with the largest fetch (the fill before it reads nothing)
    cf, cw = max(cfs), max(cws)
    kf, kw = COPY_BYTES / 1024 / cf, COPY_BYTES / 1024 / cw
    print(f"copy calibration ({COPY_BYTES} B read + written): FETCH_SIZE {cf:.1f} KiB -> factor {kf:.4f}; "
          f"WRITE_SIZE {cw:.1f} KiB -> factor {kw:.4f}")
    fe, wr = _pmc(fetch_dir, "FETCH_SIZE"), _pmc(write_dir, "WRITE_SIZE")
    print(f"{'kernel':34s} {'disp':>5s} {'live':>5s} {'fetch_MB/live':>14s} {'write_MB/live':>14s} {'total_MB/live':>14s}")
    for k in sorted(fe, key=lambda k: -sum(fe[k])):
        f = sorted(fe[k])
        w = sorted(wr.get(k, [0.0]))
        # live dispatches move at least 1 MiB; predicated-off ones read descriptors only
        fl = [x for x in f if x * kf * 1024 > (1 << 20)]
        wl = w[len(w) - len(fl):] if fl else []
        if not fl:
            print(f"{k[:34]:34s} {len(f):5d} {0:5d} {'-':>14s} {'-':>14s} {'-':>14s}")
            continue
        fb = sum(fl) / len(fl) * kf * 1024 / 1e6
        wb = sum(wl) / max(1, len(wl)) * kw * 1024 / 1e6
        print(f"{k[:34]:34s} {len(f):5d} {len(fl):5d} {fb:14.2f} {wb:14.2f} {fb + wb:14.2f}")


def copy_calibration(nbytes=1 << 30, reps=4):
    """A plain streaming copy of known size on the GPU: dst = -src over nbytes of
    FP32, reps times.  Reads and writes nbytes per launch; profiling its
    FETCH_SIZE / WRITE_SIZE gives the scale factor for the HBM traffic counters."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                                           "launches": reps}}))


def profiled_traffic(kernel, prec):
    """HBM bytes per launch of `kernel` as last measured with rocprofv3
    FETCH_SIZE + WRITE_SIZE (copy-calibrated), from the summaries in profiles/."""
    import glob
    import re
    key = re.escape(f"traffic_bytes_per_launch[{kernel},{prec}]")
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(f) as fh:
            for line in fh:
                m = re.match(key + r"\s*=\s*([0-9.eE+]+)", line.strip())
                if m:
                    return float(m.group(1))
    return None


def sweep_bytes(S, prec):
    """Algorithmic HBM bytes of one step's S live push_sweep launches, per particle.
    Sweep k < S-1 reads X, V (and the iterate v_k for k > 0) and writes v_{k+1};
    the last sweep reads X, V, v_k (mixed: also the low velocity word) and writes
    X, V (mixed: V as hi and lo words)."""
    if prec == "mixed":
        first, mid, last0, last = 12, 16, 24, 28
    else:
        first, mid, last0, last = 24, 32, 32, 40
    if S <= 1:
        return last0
    return first + mid * (S - 2) + last


def timed_run(cfg, x, v, steps, warmup, profile_steps, autotune=True):
    import picgolf
    with picgolf.Context(cfg) as c:
        c.upload(x, v)
        tune = c.autotune(2) if autotune else None
        if warmup:
            c.run(warmup)
        t0 = time.perf_counter()
        D, S = c.run(steps)          # synchronous: returns after the device finished
        t1 = time.perf_counter()
        prof = None
        if profile_steps:
            _, _, prof = c.run_profiled(profile_steps)
    return (t1 - t0), D[0], S[0], prof, tune


def bench_ensemble(steps=10, warmup=10, n=512, autotune=True):
    """The batched ensemble: n independent two-stream instances with uneven N,
    P, W and beam speed in one context, mixed and FP64.  Field-energy error of
    the mixed run against the FP64 run over the timed steps (linear phase)."""
    import numpy as np
    import picgolf
    from picgolf import configs
    res = {}
    for prec in ("mixed", "fp64"):
        cfgs, states, _ = configs.ensemble(n, precision=prec)
        with picgolf.Context(cfgs) as c:
            for i, (x, v) in enumerate(states):
                c.upload(x, v, inst=i)
            tune = c.autotune(2) if autotune else None
            c.run(warmup)
            t0 = time.perf_counter()
            D, S = c.run(steps)
            secs = time.perf_counter() - t0
        Ptot = sum(k.P for k in cfgs)
        res[prec] = dict(pushes=Ptot * steps / secs, ms=secs / steps * 1e3, D=D, S=S, Ptot=Ptot, tune=tune)
    f, m = res["fp64"], res["mixed"]
    rel = np.abs(m["D"][:, :, 0] - f["D"][:, :, 0]) / np.abs(f["D"][:, :, 0])
    return {"workload": f"ensemble of {n} two-stream instances, N in 32..1024, 16-256 particles per cell, "
                        f"Gaussian +-7 midpoint, rtol=1e-8 (sum P = {m['Ptot']})",
            "mixed_pushes_per_s": m["pushes"], "fp64_pushes_per_s": f["pushes"],
            "mixed_over_fp64": m["pushes"] / f["pushes"], "mixed_ms_per_step": m["ms"],
            "fp64_ms_per_step": f["ms"], "mean_sweeps_mixed": float(m["S"].mean()),
            "field_energy_rel_err_mixed_vs_fp64": float(rel.max()), "steps": steps, "warmup": warmup,
            "tiling": {"mixed": m["tune"], "fp64": f["tune"]}}


def bench_2d(steps=500):
    """The electrostatic 2D3V workload (256 x 256, 64 particles per cell), the
    whole 500-step run timed, mixed and FP64, with the script's area pairing
    and with the textbook (continuous) one.  The script's pairing makes the
    run chaotic in the particle positions; an FP64 run from positions
    displaced at random by 1e-8 of a cell is reported beside it as the
    control of what any inexact arithmetic can reach."""
    import numpy as np
    import picgolf
    from picgolf import configs

    def one(pairing, dx_cells=0.0, precs=("mixed", "fp64")):
        res = {}
        for prec in precs:
            cfg, st, _ = configs.es2d3v(precision=prec, pairing=pairing)
            x, y = st[0], st[1]
            if dx_cells:
                rng = np.random.default_rng(77)
                x = x + dx_cells * rng.uniform(-1, 1, x.size) / cfg.NX
                y = y + dx_cells * rng.uniform(-1, 1, y.size) / cfg.NY
                x = np.where(x > 1.0, x - 1.0, np.where(x <= 0.0, x + 1.0, x))
                y = np.where(y > 1.0, y - 1.0, np.where(y <= 0.0, y + 1.0, y))
            with picgolf.Context2D(cfg) as c:
                c.upload(x, y, *st[2:])
                c.run(2)                 # warm start: graphs captured, caches warm
                c.upload(x, y, *st[2:])
                t0 = time.perf_counter()
                K = c.run(steps)
                secs = time.perf_counter() - t0
            res[prec] = dict(pushes=cfg.P * steps / secs, ms=secs / steps * 1e3, K=K, P=cfg.P)
        return res

    def summary(res):
        f, m = res["fp64"], res["mixed"]
        rel = np.abs(m["K"][:, 0] - f["K"][:, 0]) / np.abs(f["K"][:, 0])
        return {"mixed_pushes_per_s": m["pushes"], "fp64_pushes_per_s": f["pushes"],
                "mixed_over_fp64": m["pushes"] / f["pushes"], "mixed_ms_per_step": m["ms"],
                "fp64_ms_per_step": f["ms"], "steps": steps,
                "field_energy_rel_err_mixed_vs_fp64_max": float(rel.max()),
                "field_energy_rel_err_mixed_vs_fp64_first_100_steps": float(rel[:50].max()),
                "momentum_x_drift_mixed": float(np.abs(m["K"][:, 3] - m["K"][0, 3]).max())}

    ref = one("reference")
    out = {"workload": "electrostatic 2D3V 256x256, 64 particles per cell (4194304), B0 = pi/2 along x, "
                       "dt=1/1024, reference area pairing", **summary(ref)}
    ctl = one("reference", 1e-8, precs=("fp64",))["fp64"]["K"]
    Kf = ref["fp64"]["K"]
    out["fp64_control_1e-8_cell_field_energy_rel_err_max"] = float(
        np.max(np.abs(ctl[:, 0] - Kf[:, 0]) / np.abs(Kf[:, 0])))
    out["textbook_pairing"] = summary(one("textbook"))
    return out


def bench_pic2d3v(steps=20, warmup=2, N=1024, ppc=16, shape="bspline2", precision="both", cpu=True):
    """The structured 2D3V module's electrostatic loop at the grid and
    particle count of its driver script (src/2D3V.jl: 1024 x 1024 cells, 16
    particles per cell per species, electrons and ions of mass 32, quadratic
    B-splines, both species sorted per cell once), the field gathered as the
    reference's update! leaves it; FP64, then mixed from the same state, and
    an FP64 run from positions displaced by up to four 2^-32 quanta of the
    box as the control of what the mixed position resolution can reach (the
    reference's shapes move a particle's charge by two cells when it crosses
    a cell edge).  precision "mixed" or "fp64" times that run alone (for
    profiles); cpu=False leaves out the CPU restatement's sample."""
    import numpy as np
    import picgolf
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    rng = np.random.default_rng(0)
    P = N * N * ppc
    n0 = 4 * np.pi ** 2
    sps, xyvs = [], []
    for q, m, vth in ((-1.0, 1.0, 0.05), (1.0, 32.0, 0.05 / np.sqrt(32))):
        sps.append(SpeciesConfig(P=P, shape=shape, charge=q, mass=m, weight=n0 / P))
        xyvs.append(np.vstack([rng.random(P), rng.random(P), vth * rng.standard_normal((3, P))]))
    def one(prec, states):
        cfg = Config2D3V(NX=N, NY=N, dt=1.0 / (8 * N), species=sps, B0=(np.sqrt(n0) / 4, 0.0, 0.0), ntskip=4,
                         ngskip=4, precision=prec)
        with picgolf.Context2D3V(cfg) as c:
            for k, xyv in enumerate(states):
                c.upload(k, xyv)
                c.sort(k, 1.0 / N, 1.0 / N)
            c.run(warmup)
            t0 = time.perf_counter()
            rows = c.run(steps)
            secs = time.perf_counter() - t0
        return cfg, rows, secs

    if precision != "both":
        cfg, rows, secs = one(precision, xyvs)
        return {"workload": f"structured 2D3V electrostatic loop: {N}x{N} cells, 2 species x {P} particles, "
                            f"{shape}, {precision}", "pushes_per_s": 2 * P * steps / secs,
                "ms_per_step": secs / steps * 1e3, "steps": steps, "warmup": warmup}
    cfg, rows, secs = one("fp64", xyvs)
    _, rows_m, secs_m = one("mixed", xyvs)
    for xyv in xyvs:
        for k in (0, 1):
            z = xyv[k] + rng.uniform(-4, 4, P) * 2.0 ** -32
            xyv[k] = np.where(z > 1.0, z - 1.0, np.where(z > 0.0, z, z + 1.0))
    _, rows_c, _ = one("fp64", xyvs)
    del xyvs

    def rel(a, col):
        return float(np.max(np.abs(a[:, col] - rows[:, col]) / np.abs(rows[:, col])))
    out = {"workload": f"structured 2D3V electrostatic loop: {N}x{N} cells, 2 species x {P} particles "
                       f"({ppc} per cell each), {shape}, ions of mass 32, B0 along x, dt=1/(8N), "
                       f"diagnostics every 4 steps, FP64 (mixed under 'mixed'), particles sorted per cell once "
                       f"before the timed steps",
           "pushes_per_s": 2 * P * steps / secs, "ms_per_step": secs / steps * 1e3, "steps": steps,
           "warmup": warmup, "kinetic_energy_first_last": [float(rows[0, 0]), float(rows[-1, 0])],
           "mixed": {"pushes_per_s": 2 * P * steps / secs_m, "ms_per_step": secs_m / steps * 1e3,
                     "mixed_over_fp64": secs / secs_m,
                     "field_energy_rel_err_vs_fp64": rel(rows_m, 1), "kinetic_rel_err_vs_fp64": rel(rows_m, 0),
                     "fp64_displaced_control_field_energy_rel_err": rel(rows_c, 1),
                     "fp64_displaced_control_kinetic_rel_err": rel(rows_c, 0)}}
    if not cpu:
        return out
    # the CPU restatement on a sample: the same grid, 1/32 of each species (sorted), two steps
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = host_cpu()[2]
    n = P // 32
    sample = []
    for sp in sps:
        xyv = np.vstack([rng.random(n), rng.random(n), 0.05 * rng.standard_normal((3, n))])
        xyv, _ = oracle.sort23(xyv, 1.0 / N, 1.0 / N)
        sample.append(dict(shape=shape, charge=sp.charge, mass=sp.mass, weight=n0 / n, xyv=xyv))
    t0 = time.perf_counter()
    oracle.run23(N, N, 1.0, 1.0, cfg.dt, cfg.B0, sample, 2, ntskip=4, ngskip=4, nthreads=threads)
    cpu_secs = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": 2 * n * 2 / cpu_secs, "unit": "particle-pushes/s", "cores": threads,
                           "kind": "port", "sample": f"FP64 CPU restatement, the same {N}x{N} grid and solve, "
                           f"2 species x {n} particles (1/32), 2 steps; push on {threads} threads, deposit "
                           f"and solve on one"}
    return out


def driver_2d3v_parameters(N=1024):
    """The normalised parameters src/2D3V.jl:12-70 derives (n = 1e16, B = 2.1 T,
    Te = 10 keV, kresolution = 5, L0 = 1): dt = 1/(8N), B0, n0, vth."""
    import math
    c, me, e = 299792458.0, 9.1093837e-31, 1.60217663e-19
    mu0 = 4.0e-7 * math.pi
    eps0 = 1.0 / mu0 / c / c
    n, B, M, TeeV = 1e16, 2.1, 2 * 1836, 1e4
    Va = B / math.sqrt(mu0 * M * me * n)
    Oi0 = e * B / M / me
    L = Va / Oi0 * 2 * math.pi * 5
    T = L / c
    phi_s = me * (L / T) ** 2 / e
    n_s = phi_s * eps0 / L ** 2 / e
    A_s = T * phi_s / L
    return {"dt": L / N / c / T / 8, "B0": B / (A_s / L), "n0": n / n_s, "vth": math.sqrt(TeeV * e * 2 / me) / L * T}


def bench_lorenz2d3v(steps=10, warmup=2, N=1024, ppc=16, shape="bspline2", field="lorenz_staggered"):
    """The structured 2D3V module's driver whole (src/2D3V.jl:50-127):
    LorenzGaugeStaggeredField(1024, 1024; ImEx(1), buffer = 10, B0 along x),
    electrons and ions of mass 32 with 16 particles per cell each, quadratic
    B-splines, sorted per cell once, warmup!, then loop! and diagnose! every
    step (ntskip = ngskip = 1).  FP64.  Particles are uniform random with
    Gaussian velocities of the driver's vth (not its Halton layout).
    field "lorenz_semi_implicit": the driver's commented-out alternative
    (src/2D3V.jl:104-105), LorenzGaugeSemiImplicitField with fieldimex
    ImEx(1.0), sourceimex ImEx(0.05), buffer 10, rtol sqrt(eps()),
    maxiters 1000; a push counts once per step whatever the iterations."""
    import numpy as np
    import picgolf
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    q = driver_2d3v_parameters(N)
    rng = np.random.default_rng(0)
    P = N * N * ppc
    sps, xyvs = [], []
    for ch, m in ((-1.0, 1.0), (1.0, 32.0)):
        vth = q["vth"] / np.sqrt(m)
        sps.append(SpeciesConfig(P=P, shape=shape, charge=ch, mass=m, weight=q["n0"] / P))
        xyvs.append(np.vstack([rng.random(P), rng.random(P), vth / np.sqrt(2) * rng.standard_normal((3, P))]))
    semi = field == "lorenz_semi_implicit"
    cfg = Config2D3V(NX=N, NY=N, dt=q["dt"], species=sps, B0=(q["B0"], 0.0, 0.0), buffer=10, ntskip=1, ngskip=1,
                     field=field, theta=1.0, theta_source=0.05, maxiters=1000)
    with picgolf.Context2D3V(cfg) as c:
        for k, xyv in enumerate(xyvs):
            c.upload(k, xyv)
            c.sort(k, 1.0 / N, 1.0 / N)
        c.warmup()
        c.run(warmup)
        t0 = time.perf_counter()
        rows = c.run(steps)
        secs = time.perf_counter() - t0
        iters = c.lg_iterations() if semi else None
    # the CPU restatement on a sample: the same grid and transforms, 1/32 of
    # each species (sorted), warmup! and one step
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n = P // 32
    sample = []
    for sp in sps:
        xyv = np.vstack([rng.random(n), rng.random(n), q["vth"] / np.sqrt(2 * sp.mass) *
                         rng.standard_normal((3, n))])
        xyv, _ = oracle.sort23(xyv, 1.0 / N, 1.0 / N)
        sample.append(dict(shape=shape, charge=sp.charge, mass=sp.mass, weight=q["n0"] / n, xyv=xyv))
    if semi:
        sample, F, grid = oracle.lg_si_warmup(N, N, 1.0, 1.0, cfg.dt, sample, 10)
        t0 = time.perf_counter()
        o = oracle.lg_si_run(N, N, 1.0, 1.0, cfg.dt, np.array(cfg.B0), sample, F, grid, np.zeros((6, N, N)), 1,
                             theta=1.0, theta_source=0.05, maxiters=1000, buffer=10)
        cpu_secs = time.perf_counter() - t0
        cpu_iters = int(o[6][0])
        fname = "LorenzGaugeSemiImplicitField ImEx(1.0), sourceimex ImEx(0.05), rtol sqrt(eps), maxiters 1000"
    else:
        sample, F = oracle.lg_warmup(N, N, 1.0, 1.0, cfg.dt, sample)
        t0 = time.perf_counter()
        oracle.lg_run(N, N, 1.0, 1.0, cfg.dt, np.array(cfg.B0), sample, F, 1, buffer=10)
        cpu_secs = time.perf_counter() - t0
        fname = "LorenzGaugeStaggeredField ImEx(1)"
    out = {"workload": f"structured 2D3V Lorenz-gauge driver (src/2D3V.jl): {N}x{N} cells, "
                       f"{fname}, buffer 10, 2 species x {P} particles ({ppc} per cell "
                       f"each), {shape}, ions of mass 32, B0 along x, dt=1/(8N), diagnose! every step, FP64, "
                       f"particles sorted per cell once",
           "pushes_per_s": 2 * P * steps / secs, "ms_per_step": secs / steps * 1e3, "steps": steps,
           "warmup": warmup, "kinetic_energy_first_last": [float(rows[0, 0]), float(rows[-1, 0])],
           "field_energy_first_last": [float(rows[0, 1]), float(rows[-1, 1])],
           "cpu_baseline": {"value": 2 * n / cpu_secs, "unit": "particle-pushes/s", "cores": 1, "kind": "port",
                            "sample": f"FP64 CPU restatement, the same {N}x{N} grid and transforms, 2 species x "
                            f"{n} particles (1/32), one step with diagnose!, one thread"}}
    if semi:
        out["iterations_per_step"] = [int(k) for k in iters]
        out["cpu_baseline"]["sample"] += f" ({cpu_iters} iterations)"
    return out


def em1d_setup(N, P, scheme, seed=0):
    """The scripts' set-up: eps = 1/10, dt = eps^2/N/2, W = 400, w = W/P*N,
    x = rand(P); the fixed point's v = (sin a/4, cos a/4, 0) with a = 2pi*rand,
    the explicit script's v = randn(P, 3)."""
    import numpy as np
    eps = 0.1
    dt = eps ** 2 / N / 2
    w = 400 / P * N
    rng = np.random.default_rng(seed)
    x = rng.random(P)
    if scheme == "explicit":
        v = rng.standard_normal((3, P))
    else:
        a = 2 * np.pi * rng.random(P)
        v = np.vstack([np.sin(a) / 4, np.cos(a) / 4, np.zeros(P)])
    return eps, dt, w, x, v


def bench_em1d(T=4048, scaled_steps=20, scheme="fixed_point"):
    """A 1D electromagnetic script whole (src/GaussianEMFixedPoint.jl, or
    src/GaussianEMExplicit.jl with scheme "explicit"): N=64, P=16N,
    eps=1/10, dt=eps^2/N/2, W=400, T=4048 steps with the twelve fields of
    every step and the per-step sort; and the same loop at N=1024 with 2^22
    particles.  FP64."""
    import picgolf
    from picgolf.core_em1d import ConfigEM1D
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    res = {}
    for name, N, P, steps in (("script", 64, 1024, T), ("scaled", 1024, 1 << 22, scaled_steps)):
        eps, dt, w, x, v = em1d_setup(N, P, scheme)
        with picgolf.ContextEM1D(ConfigEM1D(N=N, P=P, eps=eps, dt=dt, w=w, scheme=scheme)) as c:
            c.upload(x, v)
            c.run(2)                       # graph captured, caches warm
            c.upload(x, v)
            t0 = time.perf_counter()
            f, sw = c.run(steps)
            secs = time.perf_counter() - t0
        res[name] = {"N": N, "P": P, "steps": steps, "ms_per_step": secs / steps * 1e3,
                     "pushes_per_s": P * steps / secs, "mean_sweeps": float(sw.mean())}
    # the CPU restatement: the script's set-up, 40 steps
    N, P = 64, 1024
    eps, dt, w, x, v = em1d_setup(N, P, scheme)
    t0 = time.perf_counter()
    if scheme == "explicit":
        oracle.emx_run(N, eps, dt, w, x, v, 40)
    else:
        oracle.em_run(N, eps, dt, w, x, v, 40)
    cpu = time.perf_counter() - t0
    script = "src/GaussianEMExplicit.jl" if scheme == "explicit" else "src/GaussianEMFixedPoint.jl"
    what = "explicit step" if scheme == "explicit" else "fixed point"
    return {"workload": f"1D electromagnetic {what} ({script}): the script "
                        f"(N=64, P=1024, {T} steps, fields every step, sort every step) and the same loop at "
                        f"N=1024, P=2^22; FP64",
            **res,
            "cpu_baseline": {"value": P * 40 / cpu, "unit": "particle-pushes/s", "cores": 1, "kind": "port",
                             "sample": "FP64 CPU restatement, the script's set-up, 40 steps, one thread"}}


def host_cpu():
    """(nproc, CPU model name, threads the CPU baselines use).  The threads are
    the job's CPU share: OMP_NUM_THREADS where the host sets it (16 per GPU on
    the MI355X boxes), else every core nproc reports."""
    nproc = os.cpu_count() or 1
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip() or nproc)
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = os.environ.get("OMP_NUM_THREADS")
    threads = int(share) if share and share.isdigit() and int(share) > 0 else nproc
    return nproc, model, max(1, min(threads, nproc))


def cpu_baseline(cfg, x, v, budget_s=12.0):
    """The CPU restatement (FP64) timed on this host on a bounded sample: the
    same N and stencil with the first n particles, steps until ~budget_s, on
    the job's CPU share (OpenMP) and on one thread (the scripts' own loop is
    single-threaded)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nproc, model, threads = host_cpu()
    n = 2 ** 14
    w = cfg.W / n * cfg.N

    def timed(nthreads, budget, max_steps):
        xs, vs = x[:n].copy(), v[:n].copy()
        t0 = time.perf_counter()
        steps = 0
        E = None
        while True:
            xs, vs, E, _, _ = oracle.run(cfg.N, 7, 1, cfg.W, w, cfg.dt, xs, vs, 1, rtol=cfg.rtol, E=E,
                                         nthreads=nthreads)
            steps += 1
            if time.perf_counter() - t0 > budget or steps >= max_steps:
                break
        return n * steps / (time.perf_counter() - t0), steps

    value, steps = timed(threads, budget_s, 200)
    single, steps1 = timed(1, budget_s / 3, 50)
    return {"value": value, "unit": "particle-pushes/s", "cores": threads, "kind": "port",
            "sample": f"FP64 CPU restatement, N=4096, first {n} particles of the workload (w rescaled), "
                      f"{steps} steps, OpenMP {threads} threads (the job's CPU share; nproc {nproc})",
            "nproc": nproc, "cpu_model": model,
            "single_thread": {"value": single, "steps": steps1, "cores": 1}}


def bench_small():
    """The two small workloads whole, through both run paths: the NGP leapfrog
    two-stream run (src/NGPFourier.jl: N=128, P=8192, 1024 steps) and the quiet
    Gaussian fixed-point run (src/GaussianFixedPointQuiet.jl: N=64, P=2048,
    8192 steps, rtol=4eps), FP64; the CPU restatement on one thread beside
    them (the scripts are single-threaded), on the whole NGP run and on the
    first 512 steps of the quiet run."""
    import picgolf
    from picgolf import configs
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    out = {}
    for name, fn, shape, scheme, cpu_steps in (("ngp_two_stream", configs.ngp_two_stream, 0, 0, 1024),
                                               ("quiet_two_stream", configs.quiet_two_stream, 7, 1, 512)):
        cfg, x, v, T = fn()
        rec = {"N": cfg.N, "P": cfg.P, "steps": T}
        for path in ("graph", "persistent"):
            with picgolf.Context(cfg) as c:
                c.set_run_path(path)
                c.upload(x, v)
                c.run(min(T, 64))                 # graphs captured, code loaded
                c.upload(x, v)
                t0 = time.perf_counter()
                D, S = c.run(T)
                secs = time.perf_counter() - t0
            rec[path] = {"us_per_step": secs / T * 1e6, "pushes_per_s": cfg.P * T / secs,
                         "mean_sweeps": float(S.mean())}
        rec["persistent_over_graph"] = rec["graph"]["us_per_step"] / rec["persistent"]["us_per_step"]
        t0 = time.perf_counter()
        oracle.run(cfg.N, shape, scheme, cfg.W, cfg.w, cfg.dt, x, v, cpu_steps, rtol=cfg.rtol, nthreads=1)
        cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": cfg.P * cpu_steps / cpu, "unit": "particle-pushes/s", "cores": 1,
                               "kind": "port", "sample": f"FP64 CPU restatement, the same run, first {cpu_steps} "
                                                         f"steps, one thread"}
        out[name] = rec
    return out


def full_run_accuracy(every=50):
    """The headline workload over its whole configured run (2000 steps) in
    both precisions: the mixed path's field-energy relative error against
    FP64 at every step, its momentum drift, and the fitted rate of mode 2 of
    the field (least squares on log|E^_2| sampled every `every` steps)."""
    import numpy as np
    import picgolf
    from picgolf import configs
    out, rates = {}, {}
    for prec in ("mixed", "fp64"):
        cfg, x, v, T = configs.scaled_two_stream(precision=prec)
        Ds, a2 = [], []
        t0 = time.perf_counter()
        with picgolf.Context(cfg) as c:
            c.upload(x, v)
            for _ in range(T // every):
                D, _ = c.run(every)
                Ds.append(D[0])
                a2.append(abs(np.fft.fft(c.download_field())[2]))
        out[prec] = np.concatenate(Ds)
        t = np.arange(1, T // every + 1) * every * cfg.dt
        rates[prec] = float(np.polyfit(t, np.log(a2), 1)[0])
        print(f"[bench] full run {prec}: {T} steps in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    m, f = out["mixed"], out["fp64"]
    rel = np.abs(m[:, 0] - f[:, 0]) / np.abs(f[:, 0])
    return {"steps": int(m.shape[0]), "field_energy_rel_err_max": float(rel.max()),
            "field_energy_rel_err_final": float(rel[-1]),
            "momentum_drift_mixed": float(np.max(np.abs(m[:, 3] - m[0, 3]))),
            "mode2_rate_mixed": rates["mixed"], "mode2_rate_fp64": rates["fp64"],
            "mode2_rate_ratio": rates["mixed"] / rates["fp64"]}


def main():
    a = parse()
    if a.profile_steps < 1:
        sys.exit("--profile-steps must be >= 1 (the roofline uses the profiled push launches)")
    if a.copy_calibration:
        copy_calibration()
        return
    if a.summarize_trace:
        summarize_trace(a.summarize_trace)
        return
    if a.summarize_pmc:
        summarize_pmc(*a.summarize_pmc)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the multi-GPU sharding is a design "
              f"note, not built", file=sys.stderr)
        sys.exit(2)
    if a.workload == "small":
        print(json.dumps(bench_small()))
        return
    if a.workload == "ensemble":
        print(json.dumps(bench_ensemble(steps=a.steps, warmup=a.warmup, autotune=not a.no_autotune)))
        return
    if a.workload == "es2d3v":
        print(json.dumps(bench_2d(steps=a.steps)))
        return
    if a.workload == "pic2d3v":
        print(json.dumps(bench_pic2d3v(steps=a.steps, warmup=a.warmup, precision=a.precision,
                                       cpu=not a.no_cpu_baseline)))
        return
    if a.workload in ("lorenz2d3v", "lorenz2d3v_si"):
        f = "lorenz_semi_implicit" if a.workload == "lorenz2d3v_si" else "lorenz_staggered"
        print(json.dumps(bench_lorenz2d3v(steps=a.steps, warmup=a.warmup, field=f)))
        return
    if a.workload in ("em1d", "em1d_explicit"):
        scheme = "explicit" if a.workload == "em1d_explicit" else "fixed_point"
        print(json.dumps(bench_em1d(T=a.script_steps, scaled_steps=a.steps, scheme=scheme)))
        return
    import numpy as np
    from picgolf import configs
    precs = ["mixed", "fp64"] if a.precision == "both" else [a.precision]
    results = {}
    base = None
    for prec in precs:
        cfg, x, v, _ = configs.scaled_two_stream(precision=prec, P=a.particles)
        base = (cfg, x, v)
        secs, D, S, prof, tune = timed_run(cfg, x, v, a.steps, a.warmup, a.profile_steps, not a.no_autotune)
        results[prec] = dict(secs=secs, D=D, S=S, prof=prof, pushes=cfg.P * a.steps / secs, tune=tune)
        print(f"[bench] {prec}: {a.steps} steps in {secs:.3f}s, {results[prec]['pushes']:.4g} pushes/s, "
              f"mean sweeps {S.mean():.2f}", file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in results else "fp64"
    r = results[head]
    cfg = base[0]
    P = cfg.P
    # dominant kernel: the per-sweep push (gather + next deposit)
    launches, ms = r["prof"]["push_sweep"]
    avg_ms = ms / launches
    S = launches / a.profile_steps
    bytes_per_launch = P * sweep_bytes(round(S), head) / S
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "particle-pushes/s on 1 MI355X, mixed vs FP64 plain, at <=1e-6 field-energy error",
        "value": r["pushes"],
        "unit": "particle-pushes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["secs"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "mixed" else "f64",
        "data": "synthetic (seeded uniform positions, +-1 beams + N(0,0.05^2))",
        "config": {"workload": "scaled_two_stream: N=4096, P=2^24, Gaussian +-7 implicit midpoint, "
                               "rtol=1e-8, dt=1/6N, W=400", "N": cfg.N, "P": P,
                   "precision": head, "parallelism": "none (1 GPU)"},
        "mean_sweeps": float(r["S"].mean()),
        "tiling": {k: v["tune"] for k, v in results.items()},
        "kernel_avg_ms": {k: v[1] / v[0] for k, v in r["prof"].items()},
        "kernel_launches_per_step": {k: v[0] / a.profile_steps for k, v in r["prof"].items()},
        "roofline": {"bound": "hbm", "kernel": "push_sweep", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic("push_sweep", head), "avg_launch_ms": avg_ms,
                     "bytes_per_launch": bytes_per_launch},
    }
    if "fp64" in results and "mixed" in results:
        f, m = results["fp64"], results["mixed"]
        out["fp64_value"] = f["pushes"]
        out["mixed_over_fp64"] = m["pushes"] / f["pushes"]
        d1f, d1m = f["D"][:, 0], m["D"][:, 0]
        out["field_energy_rel_err_mixed_vs_fp64"] = float(np.max(np.abs(d1m - d1f) / np.abs(d1f)))
        out["momentum_drift_mixed"] = float(np.max(np.abs(m["D"][:, 3] - m["D"][0, 3])))
        if not a.no_full_run and a.particles == 2 ** 24:
            out["full_run"] = full_run_accuracy()
    if not a.no_extra:
        out["workloads"] = {"small": bench_small(),
                            "ensemble_512": bench_ensemble(autotune=not a.no_autotune), "es2d3v_256": bench_2d(),
                            "pic2d3v_1024": bench_pic2d3v(), "lorenz2d3v_1024": bench_lorenz2d3v(),
                            "lorenz2d3v_si_1024": bench_lorenz2d3v(field="lorenz_semi_implicit"),
                            "em1d": bench_em1d(),
                            "em1d_explicit": bench_em1d(scheme="explicit")}
        print(f"[bench] extra workloads: {json.dumps(out['workloads'])}", file=sys.stderr, flush=True)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(*base)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
