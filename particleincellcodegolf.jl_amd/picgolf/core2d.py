"""Python mirror of the 2D3V operators of src/Electrostatic2D3V.jl on top of
the C ABI in include/pic2d.h.

    g(z, NZ), eval(F1, F2, x, y)     src/Electrostatic2D3V.jl:84-103    -> gather
    boris(vx, vy, vz, Ex, Ey, dt)    :32-41                             -> boris
    gather, boris, move (unimod)     :128-134                           -> push
    depositcharge! + sum(ns, dims=3) :105-109, :139-141                 -> deposit
    field invert + field solve       :70-81, :142-157                   -> fieldsolve
    K[ti, 1:5]                       :166-170                           -> diagnostics
    for t in 1:T ... end             :120-176                           -> run

Grids are (NY, NX) numpy arrays in C order, i.e. Julia's column-major
NX x NY with x fastest.  Every failure raises PicError with the library's
message.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L

PAIRINGS = {"reference": L.PAIR_REFERENCE, "textbook": L.PAIR_TEXTBOOK}
PRECISIONS = {"fp64": L.FP64, "mixed": L.MIXED}


@dataclass
class Config2D:
    """NX x NY periodic grid on the unit square, P particles of weight w
    (n0/P/(dx*dy) in the script, :25), time step dt, B = B0 along x, and
    diagnostics every ns steps (the script's NS)."""
    NX: int
    NY: int
    P: int
    dt: float
    w: float
    B0: float
    ns: int = 2
    precision: str = "fp64"
    pairing: str = "reference"
    device: int = 0
    sort_every: int = 0

    def to_c(self):
        c = L.pic2d_config()
        c.NX, c.NY, c.P = int(self.NX), int(self.NY), int(self.P)
        c.precision = PRECISIONS[self.precision] if isinstance(self.precision, str) else int(self.precision)
        c.pairing = PAIRINGS[self.pairing] if isinstance(self.pairing, str) else int(self.pairing)
        c.dt, c.w, c.B0 = float(self.dt), float(self.w), float(self.B0)
        c.ns, c.device, c.sort_every = int(self.ns), int(self.device), int(self.sort_every)
        return c


class Context2D:
    def __init__(self, cfg: Config2D):
        self._lib = L.load()
        self.cfg = cfg
        h = ctypes.c_void_p()
        c = cfg.to_c()
        rc = self._lib.pic2d_create(ctypes.byref(h), ctypes.byref(c))
        if rc != 0:
            raise L.PicError(rc, f"pic2d_create: {self._lib.pic2d_last_error(None).decode()}")
        self._h = h
        self.steps = 0

    def _check(self, rc, what):
        if rc != 0:
            raise L.PicError(rc, f"{what}: {self._lib.pic2d_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pic2d_destroy(self._h)
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

    def _grid(self):
        return np.zeros((self.cfg.NY, self.cfg.NX))

    # -- device-resident runs ----------------------------------------------
    def upload(self, x, y, vx, vy, vz, Ex=None, Ey=None):
        arrs = [L.f64(a) for a in (x, y, vx, vy, vz)]
        if any(a.size != self.cfg.P for a in arrs):
            raise ValueError(f"expected {self.cfg.P} particles")
        Ex = None if Ex is None else L.f64(Ex)
        Ey = None if Ey is None else L.f64(Ey)
        self._check(self._lib.pic2d_upload(self._h, *[L.ptr(a) for a in arrs], L.ptr(Ex), L.ptr(Ey)),
                    "pic2d_upload")
        self.steps = 0

    def download(self):
        P = self.cfg.P
        out = [np.zeros(P) for _ in range(5)] + [self._grid(), self._grid()]
        self._check(self._lib.pic2d_download(self._h, *[L.ptr(a) for a in out]), "pic2d_download")
        return tuple(out)

    def run(self, nsteps):
        """nsteps of the loop; returns K with one row per step t % ns == 0."""
        ns = self.cfg.ns
        rows = (self.steps + nsteps) // ns - self.steps // ns
        K = np.zeros((rows, 5))
        self._check(self._lib.pic2d_run(self._h, int(nsteps), L.ptr(K) if rows else None), "pic2d_run")
        self.steps += nsteps
        return K

    # -- operators (FP64 contexts) -------------------------------------------
    def deposit(self, x, y, w=None):
        x, y = L.f64(x), L.f64(y)
        w = self.cfg.w if w is None else w
        rho = self._grid()
        self._check(self._lib.pic2d_deposit(self._h, L.ptr(x), L.ptr(y), x.size, float(w), L.ptr(rho)),
                    "pic2d_deposit")
        return rho

    def fieldsolve(self, rho, want_phi=False):
        rho = L.f64(rho)
        Ex, Ey = self._grid(), self._grid()
        phi = self._grid() if want_phi else None
        self._check(self._lib.pic2d_fieldsolve(self._h, L.ptr(rho), L.ptr(Ex), L.ptr(Ey), L.ptr(phi)),
                    "pic2d_fieldsolve")
        return (Ex, Ey, phi) if want_phi else (Ex, Ey)

    def gather(self, Ex, Ey, x, y):
        Ex, Ey, x, y = L.f64(Ex), L.f64(Ey), L.f64(x), L.f64(y)
        ex, ey = np.zeros(x.size), np.zeros(x.size)
        self._check(self._lib.pic2d_gather(self._h, L.ptr(Ex), L.ptr(Ey), L.ptr(x), L.ptr(y), x.size,
                                           L.ptr(ex), L.ptr(ey)), "pic2d_gather")
        return ex, ey

    def boris(self, vx, vy, vz, ex, ey, dt=None, B0=None):
        """In place on vx, vy, vz (numpy float64 arrays)."""
        dt = self.cfg.dt if dt is None else dt
        B0 = self.cfg.B0 if B0 is None else B0
        a, b, c = (L.f64(v).copy() for v in (vx, vy, vz))
        ex, ey = L.f64(ex), L.f64(ey)
        self._check(self._lib.pic2d_boris(self._h, L.ptr(a), L.ptr(b), L.ptr(c), L.ptr(ex), L.ptr(ey), a.size,
                                          float(dt), float(B0)), "pic2d_boris")
        vx[...], vy[...], vz[...] = a, b, c
        return vx, vy, vz

    def push(self, x, y, vx, vy, vz, Ex, Ey, dt=None, B0=None):
        """Gather, Boris and move, in place on the five particle arrays."""
        dt = self.cfg.dt if dt is None else dt
        B0 = self.cfg.B0 if B0 is None else B0
        arrs = [L.f64(a).copy() for a in (x, y, vx, vy, vz)]
        Ex, Ey = L.f64(Ex), L.f64(Ey)
        self._check(self._lib.pic2d_push(self._h, *[L.ptr(a) for a in arrs], arrs[0].size, L.ptr(Ex), L.ptr(Ey),
                                         float(dt), float(B0)), "pic2d_push")
        for dst, src in zip((x, y, vx, vy, vz), arrs):
            dst[...] = src
        return x, y, vx, vy, vz

    def diagnostics(self, Ex, Ey, vx, vy, w=None):
        Ex, Ey, vx, vy = L.f64(Ex), L.f64(Ey), L.f64(vx), L.f64(vy)
        w = self.cfg.w if w is None else w
        K = np.zeros(5)
        self._check(self._lib.pic2d_diagnostics(self._h, L.ptr(Ex), L.ptr(Ey), L.ptr(vx), L.ptr(vy), vx.size,
                                                float(w), L.ptr(K)), "pic2d_diagnostics")
        return K
