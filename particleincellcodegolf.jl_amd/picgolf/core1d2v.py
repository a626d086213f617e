"""Python mirror of the magnetised 1D2V scripts (src/NGP1D2V.jl,
src/NGP1D2V2S.jl) on top of the C ABI in include/pic1d2v.h.

    rho() / rho(x)                    src/NGP1D2V2S.jl:28-29, NGP1D2V.jl:38   -> deposit
    sum(k->E[k[1]]*k[2], d(x[j]))     src/NGP1D2V2S.jl:35                     -> gather
    boris(vx, vy, E, B, dt, q_m)      src/NGP1D2V2S.jl:5-11                   -> boris
    for t in 1:T ... end              src/NGP1D2V2S.jl:31-57                  -> run, fields

run() and fields() return what the scripts return: D and Es divided by the
bin length (T/TO) after the loop.  Every failure raises PicError with the
library's message.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

SHAPES = {"gauss6": L.SHAPE_GAUSS6, "gauss7": L.SHAPE_GAUSS7}


@dataclass
class Config1D2V:
    """N cells, P particles per species, time bins of `bin` steps (the
    scripts' T/TO), per-species charge q, charge-to-mass ratio q_m and kinetic
    weight mass, particle weight w, n0, B0 along z and dt; precision "fp64"
    (bitwise equal to the CPU restatement) or "mixed"."""
    N: int
    P: int
    bin: int
    w: float
    n0: float
    B0: float
    dt: float
    q: tuple = (1.0,)
    q_m: tuple = (1.0,)
    mass: tuple = (1.0,)
    shape: str = "gauss7"
    device: int = 0
    precision: str = "fp64"

    @property
    def nspecies(self):
        return len(self.q)

    def to_c(self):
        c = L.pic1d2v_config()
        c.N, c.P, c.bin = int(self.N), int(self.P), int(self.bin)
        c.shape = SHAPES[self.shape] if isinstance(self.shape, str) else int(self.shape)
        c.nspecies = self.nspecies
        for s in range(self.nspecies):
            c.q[s], c.q_m[s], c.mass[s] = float(self.q[s]), float(self.q_m[s]), float(self.mass[s])
        c.w, c.n0, c.B0, c.dt = float(self.w), float(self.n0), float(self.B0), float(self.dt)
        c.device = int(self.device)
        c.precision = {"fp64": L.FP64, "mixed": L.MIXED}[self.precision] if isinstance(self.precision, str) \
            else int(self.precision)
        return c


class Context1D2V:
    def __init__(self, cfg: Config1D2V):
        self._lib = L.load()
        self.cfg = cfg
        h = ctypes.c_void_p()
        c = cfg.to_c()
        rc = self._lib.pic1d2v_create(ctypes.byref(h), ctypes.byref(c))
        if rc != 0:
            raise L.PicError(rc, f"pic1d2v_create: {self._lib.pic1d2v_last_error(None).decode()}")
        self._h = h
        self.steps = 0

    def _check(self, rc, what):
        if rc != 0:
            raise L.PicError(rc, f"{what}: {self._lib.pic1d2v_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pic1d2v_destroy(self._h)
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

    def upload(self, species, x, vx, vy):
        arrs = [L.f64(a) for a in (x, vx, vy)]
        if any(a.size != self.cfg.P for a in arrs):
            raise ValueError(f"expected {self.cfg.P} particles")
        self._check(self._lib.pic1d2v_upload(self._h, int(species), *[L.ptr(a) for a in arrs]), "pic1d2v_upload")
        self.steps = 0

    def download(self, species):
        out = [np.zeros(self.cfg.P) for _ in range(3)]
        self._check(self._lib.pic1d2v_download(self._h, int(species), *[L.ptr(a) for a in out]), "pic1d2v_download")
        return tuple(out)

    def run_raw(self, nsteps):
        """nsteps steps; the D rows of the bins completed in this call, as the
        loop computes them (before the scripts' division by bin)."""
        b = self.cfg.bin
        rows = (self.steps + nsteps) // b - self.steps // b
        D = np.zeros((rows, 5))
        self._check(self._lib.pic1d2v_run(self._h, int(nsteps), L.ptr(D) if rows else None), "pic1d2v_run")
        self.steps += nsteps
        return D

    def run(self, nsteps):
        """As run_raw, divided by the bin length like the scripts' D ./= T÷TO."""
        return self.run_raw(nsteps) / self.cfg.bin

    def fields_raw(self, nbins=None):
        """Es sums per bin, (nbins, N), before the scripts' division by bin."""
        b = self.cfg.bin
        nbins = (self.steps + b - 1) // b if nbins is None else nbins
        Es = np.zeros((nbins, self.cfg.N))
        if nbins:
            self._check(self._lib.pic1d2v_fields(self._h, L.ptr(Es), int(nbins)), "pic1d2v_fields")
        return Es

    def fields(self, nbins=None):
        """The scripts' Es (./= T÷TO), transposed to (nbins, N)."""
        return self.fields_raw(nbins) / self.cfg.bin

    # -- operators -------------------------------------------------------------
    def deposit(self, xs):
        """rho() over the species position arrays xs (P each)."""
        x = np.concatenate([L.f64(a) for a in xs])
        if x.size != self.cfg.nspecies * self.cfg.P:
            raise ValueError("expected nspecies arrays of P positions")
        rho = np.zeros(self.cfg.N)
        self._check(self._lib.pic1d2v_deposit(self._h, L.ptr(x), L.ptr(rho)), "pic1d2v_deposit")
        return rho

    def gather(self, E, x):
        E, x = L.f64(E), L.f64(x)
        out = np.zeros(x.size)
        self._check(self._lib.pic1d2v_gather(self._h, L.ptr(E), L.ptr(x), x.size, L.ptr(out)), "pic1d2v_gather")
        return out

    def boris(self, vx, vy, Ej, B=None, dt=None, q_m=1.0):
        """In place on vx, vy (numpy float64 arrays)."""
        a, b = L.f64(vx).copy(), L.f64(vy).copy()
        Ej = L.f64(Ej)
        B = self.cfg.B0 if B is None else B
        dt = self.cfg.dt if dt is None else dt
        self._check(self._lib.pic1d2v_boris(self._h, L.ptr(a), L.ptr(b), L.ptr(Ej), a.size, float(B), float(dt),
                                            float(q_m)), "pic1d2v_boris")
        vx[...], vy[...] = a, b
        return vx, vy
