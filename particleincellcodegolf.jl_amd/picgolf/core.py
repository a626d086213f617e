"""Python mirror of the reference's PIC operators on top of the C ABI.

The scripts' closures map one to one onto Context methods:

    rho(x, y)                      src/GaussianFixedPointQuiet.jl:7   -> deposit
    real(ifft(fft(rho)./ik))       src/GaussianFixedPointQuiet.jl:9   -> fieldsolve
    u()                            src/NGPFourier.jl:2                -> drift_half
    u(); v += E[f.(x)]*dt          src/NGPFourier.jl:6                -> push_leapfrog
    x .= X .+ (v .+ V)/2*dt        src/GaussianFixedPointQuiet.jl:9   -> midpoint_position
    v[j] = V[j] + sum(...)*dt      src/GaussianFixedPointQuiet.jl:10  -> push_midpoint
    x .= mod.(x, 1)                src/GaussianFixedPointQuiet.jl:10  -> wrap
    isapprox(F, E, rtol=l)         src/GaussianFixedPointQuiet.jl:8   -> isapprox
    D[t, 1:4]                      src/GaussianFixedPointQuiet.jl:11  -> diagnostics
    @gif for t in 1:T ... end      src/GaussianFixedPointQuiet.jl:8   -> run

Methods that the scripts express as in-place updates (x .= ..., v[j] = ...)
update the numpy arrays passed in.  Every failure raises PicError carrying
the library's message.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

SHAPES = {"ngp": L.SHAPE_NGP, "area": L.SHAPE_AREA, "gauss6": L.SHAPE_GAUSS6, "gauss7": L.SHAPE_GAUSS7}
SCHEMES = {"leapfrog": L.LEAPFROG, "midpoint": L.MIDPOINT, "simpson": L.SIMPSON}
PRECISIONS = {"fp64": L.FP64, "mixed": L.MIXED}


def _code(v, table, what):
    if isinstance(v, str):
        if v not in table:
            raise ValueError(f"unknown {what} {v!r}; choose from {sorted(table)}")
        return table[v]
    return int(v)


@dataclass
class Config:
    """One instance.  Field meanings follow the scripts' literals
    (src/GaussianFixedPointQuiet.jl:1-5): N cells, P particles, W, w = W/P*N,
    dt, rtol and at most max_sweeps Picard sweeps per step."""
    N: int
    P: int
    W: float
    dt: float
    w: float = None
    shape: str = "gauss7"
    scheme: str = "midpoint"
    precision: str = "fp64"
    rtol: float = 1e-8
    max_sweeps: int = 10
    device: int = 0

    def __post_init__(self):
        if self.w is None:
            self.w = self.W / self.P * self.N

    def to_c(self):
        c = L.pic_config()
        c.N, c.P = int(self.N), int(self.P)
        c.shape = _code(self.shape, SHAPES, "shape")
        c.scheme = _code(self.scheme, SCHEMES, "scheme")
        c.precision = _code(self.precision, PRECISIONS, "precision")
        c.W, c.w, c.dt, c.rtol = float(self.W), float(self.w), float(self.dt), float(self.rtol)
        c.max_sweeps, c.device = int(self.max_sweeps), int(self.device)
        return c


class Context:
    """A device context holding one instance or a batch of instances."""

    def __init__(self, configs):
        self._lib = L.load()
        if isinstance(configs, Config):
            configs = [configs]
        self.configs = list(configs)
        arr = (L.pic_config * len(self.configs))(*[c.to_c() for c in self.configs])
        h = ctypes.c_void_p()
        rc = self._lib.pic_create_batch(ctypes.byref(h), arr, len(self.configs))
        if rc != 0:
            msg = self._lib.pic_last_error(None).decode()
            raise L.PicError(rc, f"pic_create_batch: {msg}")
        self._h = h
        self.N = self.configs[0].N

    # -- plumbing -------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.pic_last_error(self._h).decode()
            raise L.PicError(rc, f"{what}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pic_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- operators ------------------------------------------------------
    def deposit(self, x, y=None, w=None, shape=None):
        x = L.f64(x)
        y = None if y is None else L.f64(y)
        cfg = self.configs[0]
        w = cfg.w if w is None else w
        shape = _code(cfg.shape if shape is None else shape, SHAPES, "shape")
        rho = np.zeros(self.N)
        self._check(self._lib.pic_deposit(self._h, L.ptr(x), L.ptr(y), x.size, float(w), shape,
                                          L.ptr(rho)), "pic_deposit")
        return rho

    def fieldsolve(self, rho):
        rho = L.f64(rho)
        E = np.zeros(self.N)
        self._check(self._lib.pic_fieldsolve(self._h, L.ptr(rho), L.ptr(E)), "pic_fieldsolve")
        return E

    def drift_half(self, x, v, dt):
        xx = L.f64(x).copy()
        v = L.f64(v)
        self._check(self._lib.pic_drift_half(self._h, L.ptr(xx), L.ptr(v), xx.size, float(dt)),
                    "pic_drift_half")
        x[...] = xx
        return x

    def push_leapfrog(self, x, v, E, dt, shape=None):
        xx, vv, E = L.f64(x).copy(), L.f64(v).copy(), L.f64(E)
        shape = _code(self.configs[0].shape if shape is None else shape, SHAPES, "shape")
        self._check(self._lib.pic_push_leapfrog(self._h, L.ptr(xx), L.ptr(vv), L.ptr(E), xx.size,
                                                float(dt), shape), "pic_push_leapfrog")
        x[...] = xx
        v[...] = vv
        return x, v

    def midpoint_position(self, X, V, v, dt):
        X, V, v = L.f64(X), L.f64(V), L.f64(v)
        x = np.zeros_like(X)
        self._check(self._lib.pic_midpoint_position(self._h, L.ptr(X), L.ptr(V), L.ptr(v), L.ptr(x),
                                                    X.size, float(dt)), "pic_midpoint_position")
        return x

    def push_midpoint(self, X, V, x, v, E, dt, shape=None):
        X, V, x, E = L.f64(X), L.f64(V), L.f64(x), L.f64(E)
        vv = np.zeros_like(X)
        shape = _code(self.configs[0].shape if shape is None else shape, SHAPES, "shape")
        self._check(self._lib.pic_push_midpoint(self._h, L.ptr(X), L.ptr(V), L.ptr(x), L.ptr(vv),
                                                L.ptr(E), X.size, float(dt), shape), "pic_push_midpoint")
        v[...] = vv
        return v

    def simpson_kick(self, a, b, v_in, E, coef, shape=None):
        """v_in + sum(k->E[k[1]]*k[2], d((a+b)/2))*coef/6, one stage kick of the
        Simpson-1/3 sweep (src/GaussianFixedPointQuietSimpson13.jl:12,14,16 with
        (a, b, coef) = (X, X, dt), (X, x, 4dt), (x, x, dt)); returns v_out."""
        a, b, v_in, E = L.f64(a), L.f64(b), L.f64(v_in), L.f64(E)
        out = np.zeros_like(a)
        shape = _code(self.configs[0].shape if shape is None else shape, SHAPES, "shape")
        self._check(self._lib.pic_simpson_kick(self._h, L.ptr(a), L.ptr(b), L.ptr(v_in), L.ptr(out), L.ptr(E),
                                               a.size, float(coef), shape), "pic_simpson_kick")
        return out

    def wrap(self, x):
        xx = L.f64(x).copy()
        self._check(self._lib.pic_wrap(self._h, L.ptr(xx), xx.size), "pic_wrap")
        x[...] = xx
        return x

    def isapprox(self, F, E, rtol, atol=0.0):
        F, E = L.f64(F), L.f64(E)
        out = ctypes.c_int32()
        self._check(self._lib.pic_isapprox(self._h, L.ptr(F), L.ptr(E), F.size, float(rtol), float(atol),
                                           ctypes.byref(out)), "pic_isapprox")
        return bool(out.value)

    def diagnostics(self, E, v, W=None):
        E, v = L.f64(E), L.f64(v)
        W = self.configs[0].W if W is None else W
        out = np.zeros(4)
        self._check(self._lib.pic_diagnostics(self._h, L.ptr(E), L.ptr(v), v.size, float(W), L.ptr(out)),
                    "pic_diagnostics")
        return out

    # -- device-resident runs --------------------------------------------
    def upload(self, x, v, E=None, inst=0):
        x, v = L.f64(x), L.f64(v)
        E = None if E is None else L.f64(E)
        P = self.configs[inst].P
        if x.size != P or v.size != P:
            raise ValueError(f"instance {inst} has P={P} particles, got {x.size}/{v.size}")
        self._check(self._lib.pic_upload(self._h, inst, L.ptr(x), L.ptr(v), L.ptr(E)), "pic_upload")

    def download(self, inst=0):
        cfg = self.configs[inst]
        x, v, E = np.zeros(cfg.P), np.zeros(cfg.P), np.zeros(cfg.N)
        self._check(self._lib.pic_download(self._h, inst, L.ptr(x), L.ptr(v), L.ptr(E)), "pic_download")
        return x, v, E

    def download_field(self, inst=0):
        """The field E carried between steps, without the particles."""
        E = np.zeros(self.configs[inst].N)
        self._check(self._lib.pic_download(self._h, inst, None, None, L.ptr(E)), "pic_download")
        return E

    def run(self, nsteps, want_diag=True):
        n = len(self.configs)
        diag = np.zeros((n, nsteps, 4)) if want_diag else None
        sweeps = np.zeros((n, nsteps), dtype=np.int32) if want_diag else None
        self._check(self._lib.pic_run(self._h, int(nsteps), L.ptr(diag), L.ptr(sweeps)), "pic_run")
        return diag, sweeps

    def run_async(self, nsteps, want_diag=True):
        """Starts nsteps steps and returns at once; sync() waits for them and
        returns (diag, sweeps).  Other calls raise until then.  The arrays the
        run writes into are held here until sync(), and only a call that
        started a run replaces them."""
        n = len(self.configs)
        D = np.zeros((n, nsteps, 4)) if want_diag else None
        S = np.zeros((n, nsteps), dtype=np.int32) if want_diag else None
        self._check(self._lib.pic_run_async(self._h, int(nsteps), L.ptr(D), L.ptr(S)), "pic_run_async")
        self._pending = (D, S)

    def sync(self):
        """Waits for the run_async run; (diag, sweeps), or (None, None) when
        none was started."""
        try:
            self._check(self._lib.pic_sync(self._h), "pic_sync")
        finally:
            out, self._pending = getattr(self, "_pending", None) or (None, None), None
        return out

    # -- run path ------------------------------------------------------------
    RUN_PATHS = {"auto": L.RUN_AUTO, "graph": L.RUN_GRAPH, "persistent": L.RUN_PERSISTENT}

    def set_run_path(self, path="auto", workgroups=0):
        """How run() advances: "graph" (captured step graphs), "persistent"
        (the whole call in one launch; one small FP64 instance) or "auto"."""
        self._check(self._lib.pic_set_run_path(self._h, _code(path, self.RUN_PATHS, "run path"), int(workgroups)),
                    "pic_set_run_path")
        return self.run_path()

    def run_path(self):
        """(path the next run() takes, "graph" or "persistent"; its workgroups)."""
        p, k = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._lib.pic_get_run_path(self._h, ctypes.byref(p), ctypes.byref(k)), "pic_get_run_path")
        return ("persistent" if p.value == L.RUN_PERSISTENT else "graph"), k.value

    # -- particle tiling of the push launches ---------------------------------
    def tiling(self):
        """(particles per tile, threads per push workgroup, tiles)."""
        t, b, n = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        self._check(self._lib.pic_get_tiling(self._h, ctypes.byref(t), ctypes.byref(b), ctypes.byref(n)),
                    "pic_get_tiling")
        return t.value, b.value, n.value

    def set_tiling(self, tile_particles, block=1024):
        self._check(self._lib.pic_set_tiling(self._h, int(tile_particles), int(block)), "pic_set_tiling")
        return self.tiling()

    def autotune(self, trial_steps=2):
        """Times trial_steps steps of the current state under each candidate
        tiling, keeps the fastest and restores the state; returns the result
        as a dict."""
        r = L.pic_tuning()
        self._check(self._lib.pic_autotune(self._h, int(trial_steps), ctypes.byref(r)), "pic_autotune")
        return {"tile_particles": r.tile_particles, "block": r.block, "n_tiles": r.n_tiles,
                "candidates": r.candidates, "ms_per_step": r.ms_per_step,
                "entry_ms_per_step": r.entry_ms_per_step}

    def run_profiled(self, nsteps):
        n = len(self.configs)
        diag = np.zeros((n, nsteps, 4))
        sweeps = np.zeros((n, nsteps), dtype=np.int32)
        stats = (L.pic_kernel_stat * 16)()
        ns = ctypes.c_int32(16)
        self._check(self._lib.pic_run_profiled(self._h, int(nsteps), L.ptr(diag), L.ptr(sweeps), stats,
                                               ctypes.byref(ns)), "pic_run_profiled")
        out = {stats[i].name.decode(): (int(stats[i].launches), float(stats[i].total_ms)) for i in range(ns.value)}
        return diag, sweeps, out
