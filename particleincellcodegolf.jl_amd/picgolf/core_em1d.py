"""Python mirror of the 1D electromagnetic scripts on top of the C ABI in
include/picem1d.h: the fixed point (src/GaussianEMFixedPoint.jl, scheme
"fixed_point")

    d(y, u)                                   :11        -> deposit
    one sweep's Ex, Ey, Ez, By, Bz updates    :20-27     -> fields
    v[j,:] .= u .+ (a .+ cross(...))*dt       :28-29     -> push
    the `@gif for t ∈ 1:T` loop and `fields`  :13-50     -> upload, run, download

and the explicit step (src/GaussianEMExplicit.jl, scheme "explicit")

    Gauss, Faraday (:13, :15-16), Ampere (:23-24)        -> explicit_fields
    v += a*dt/2; v += cross(v, b)*dt; v += a*dt/2 :17-22 -> explicit_push
    the `for t ∈ 1:T` loop and `fields`       :12-30     -> upload, run, download

Velocities are a (3, P) array with rows v[:,1], v[:,2], v[:,3].  run()
returns the script's fields[:, t, 1:12] for the steps of the call as an
array of shape (12, nsteps, N) (fields[k, t] is the script's
fields[:, t, k+1]) and the sweeps of each step.  Every failure raises
PicError with the library's message.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class ConfigEM1D:
    """N cells, P particles, eps = 1/c, dt, weight w (W/P*N), the fixed
    point's rtol and sweeps, whether to sort by x after every step, and the
    scheme: "fixed_point" or "explicit"."""
    N: int
    P: int
    eps: float
    dt: float
    w: float
    rtol: float = 1e-8
    max_sweeps: int = 3
    sort: bool = True
    device: int = 0
    scheme: str = "fixed_point"

    def to_c(self):
        schemes = {"fixed_point": L.PICEM1D_FIXED_POINT, "explicit": L.PICEM1D_EXPLICIT}
        if self.scheme not in schemes:
            raise ValueError(f"scheme must be 'fixed_point' or 'explicit' (got {self.scheme!r})")
        c = L.picem1d_config()
        c.scheme = schemes[self.scheme]
        c.N, c.max_sweeps, c.P = int(self.N), int(self.max_sweeps), int(self.P)
        c.eps, c.dt, c.w, c.rtol = float(self.eps), float(self.dt), float(self.w), float(self.rtol)
        c.sort, c.device = int(bool(self.sort)), int(self.device)
        return c


class ContextEM1D:
    def __init__(self, cfg: ConfigEM1D):
        self._lib = L.load()
        self.cfg = cfg
        h = ctypes.c_void_p()
        c = cfg.to_c()
        rc = self._lib.picem1d_create(ctypes.byref(h), ctypes.byref(c))
        if rc != 0:
            raise L.PicError(rc, f"picem1d_create: {self._lib.picem1d_last_error(None).decode()}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            raise L.PicError(rc, f"{what}: {self._lib.picem1d_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.picem1d_destroy(self._h)
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

    def upload(self, x, v, E=None, B=None):
        """x (P), v (3, P), E (3, N): Ex, Ey, Ez (zeros), B (2, N): By, Bz
        (the script's zeros and ones)."""
        x, v = L.f64(x), L.f64(v).reshape(3, -1)
        Ed = None if E is None else L.f64(E).reshape(3, -1)
        Bd = None if B is None else L.f64(B).reshape(2, -1)
        self._check(self._lib.picem1d_upload(self._h, L.ptr(x), L.ptr(np.ascontiguousarray(v)), L.ptr(Ed), L.ptr(Bd)),
                    "picem1d_upload")

    def download(self):
        """(x, v (3, P), E (3, N), B (2, N))."""
        P, N = self.cfg.P, self.cfg.N
        x, v, E, B = np.zeros(P), np.zeros((3, P)), np.zeros((3, N)), np.zeros((2, N))
        self._check(self._lib.picem1d_download(self._h, L.ptr(x), L.ptr(v), L.ptr(E), L.ptr(B)), "picem1d_download")
        return x, v, E, B

    def run(self, nsteps, want_fields=True):
        N = self.cfg.N
        f = np.zeros((12, nsteps, N)) if want_fields else None
        sw = np.zeros(max(nsteps, 1), dtype=np.int32)
        self._check(self._lib.picem1d_run(self._h, int(nsteps), L.ptr(f), L.ptr(sw)), "picem1d_run")
        return f, sw[:nsteps]

    def deposit(self, y, w, u=None):
        y = L.f64(y)
        uu = None if u is None else L.f64(u)
        rho = np.zeros(self.cfg.N)
        self._check(self._lib.picem1d_deposit(self._h, L.ptr(y), L.ptr(uu), y.size, float(w), L.ptr(rho)),
                    "picem1d_deposit")
        return rho

    def fields(self, rho, Jy, Jz, Fx, Fy, Fz, Cy, Cz, E, B):
        """One sweep's updates: returns (E (3, N), B (2, N), e3 (3, N), b2 (2, N))."""
        N = self.cfg.N
        a = [L.f64(z) for z in (rho, Jy, Jz, Fx, Fy, Fz, Cy, Cz)]
        E = L.f64(E).reshape(3, N).copy()
        B = L.f64(B).reshape(2, N).copy()
        e3, b2 = np.zeros((3, N)), np.zeros((2, N))
        self._check(self._lib.picem1d_fields(self._h, *(L.ptr(z) for z in a), L.ptr(E), L.ptr(B), L.ptr(e3),
                                             L.ptr(b2)), "picem1d_fields")
        return E, B, e3, b2

    def push(self, c, e3, b2, u):
        """v (3, n) = u + (a + cross(u + a*dt/2, b))*dt at midpoints c."""
        c = L.f64(c)
        u = np.ascontiguousarray(L.f64(u).reshape(3, -1))
        v = np.zeros_like(u)
        self._check(self._lib.picem1d_push(self._h, L.ptr(c), L.ptr(np.ascontiguousarray(L.f64(e3))),
                                           L.ptr(np.ascontiguousarray(L.f64(b2))), L.ptr(u), c.size, L.ptr(v)),
                    "picem1d_push")
        return v

    def explicit_fields(self, rho, Jy, Jz, E, B):
        """The explicit step's field updates from its rho and currents:
        returns (E (3, N), B (2, N), b2 (2, N): by, bz)."""
        N = self.cfg.N
        a = [L.f64(z) for z in (rho, Jy, Jz)]
        E = L.f64(E).reshape(3, N).copy()
        B = L.f64(B).reshape(2, N).copy()
        b2 = np.zeros((2, N))
        self._check(self._lib.picem1d_explicit_fields(self._h, *(L.ptr(z) for z in a), L.ptr(E), L.ptr(B),
                                                      L.ptr(b2)), "picem1d_explicit_fields")
        return E, B, b2

    def explicit_push(self, x, E, b2, u):
        """v (3, n): the split kick of u at positions x with E (3, N) and
        b2 (2, N)."""
        x = L.f64(x)
        u = np.ascontiguousarray(L.f64(u).reshape(3, -1))
        v = np.zeros_like(u)
        self._check(self._lib.picem1d_explicit_push(self._h, L.ptr(x), L.ptr(np.ascontiguousarray(L.f64(E))),
                                                    L.ptr(np.ascontiguousarray(L.f64(b2))), L.ptr(u), x.size,
                                                    L.ptr(v)), "picem1d_explicit_push")
        return v
