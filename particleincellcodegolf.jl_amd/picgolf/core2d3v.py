"""Python mirror of the structured 2D3V module (src/PIC2D3V.jl), its
electrostatic and Lorenz-gauge paths, on top of the C ABI in
include/pic2d3v.h.

    Species(P, ...; charge, mass) / xyv    src/PIC2D3V.jl:148-208   -> SpeciesConfig, upload, download
    sort!(species, dx, dy)                 :210-215                 -> sort
    ElectrostaticField(NX, NY, Lx, Ly; dt, B0x, B0y, B0z, buffer)
                                           :270-292                 -> Config2D3V
    ElectrostaticDiagnostics(NX, NY, NT, ntskip, ngskip)
                                           :74-101                  -> Config2D3V.ntskip / ngskip
    loop!(plasma, field, to, t, _); diagnose!(diagnostics, field, plasma, t, to)
                                           :530-581, :1301-1330     -> run
    d.Exs, d.Eys, d.phis                   :1323-1327               -> fields
    deposit!, (field)(shape, x, y), boris  :1249, :1217, :49        -> deposit, gather, boris
    LorenzGaugeStaggeredField(NX, NY, Lx, Ly; dt, B0x, B0y, B0z, imex, buffer)
                                           :347-392                 -> Config2D3V(field="lorenz_staggered", theta)
    LorenzGaugeField(...; imex, buffer)    :307-345                 -> Config2D3V(field="lorenz", theta)
    LorenzGaugeSemiImplicitField(...; fieldimex, sourceimex, buffer, rtol, maxiters)
                                           :403-461                 -> Config2D3V(field="lorenz_semi_implicit",
                                                                        theta, theta_source, rtol, maxiters)
    its rhoJs+ and iteration counts        :1000-1012               -> lg_grid, lg_set_grid, lg_iterations
    warmup!(field, plasma, to)             :826-859                 -> warmup
    loop! alone (no diagnose!)             :875-979                 -> loop
    the field's phi, A, rho, J, E, B       :347-378                 -> lg_get, lg_set
    LorenzGaugeDiagnostics' Exs .. Jzs     :104-128, :1370-1386     -> lg_bins

Particles of one species are the reference's xyv: a (5, P) array with rows
x, y, vx, vy, vz (Julia's 5 x P matrix; the C side takes it column-major).
Grids are (NY, NX) numpy arrays in C order, Julia's NX x NY with x fastest.
Every failure raises PicError with the library's message.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib as L

SHAPES = {"ngp": L.NGP_2D3V, "area": L.AREA_2D3V, **{f"bspline{n}": L.BSPLINE0_2D3V + n for n in range(6)}}
FIELD_UPDATES = {"accumulate": L.FIELD_ACCUMULATE, "replace": L.FIELD_REPLACE}
BORIS = {"reference": L.BORIS_REFERENCE, "textbook": L.BORIS_TEXTBOOK}
FIELDS = {"electrostatic": L.ELECTROSTATIC_2D3V, "lorenz_staggered": L.LORENZ_STAGGERED_2D3V,
          "lorenz": L.LORENZ_2D3V, "lorenz_semi_implicit": L.LORENZ_SEMI_IMPLICIT_2D3V}


def _code(table, v):
    return table[v] if isinstance(v, str) else int(v)


@dataclass
class SpeciesConfig:
    """One Species: P particles with charge, mass, weight and shape."""
    P: int
    shape: str = "bspline2"
    charge: float = 1.0
    mass: float = 1.0
    weight: float = 1.0


@dataclass
class Config2D3V:
    """An ElectrostaticField (NX x NY cells in an Lx x Ly box, uniform B0,
    dt, ghost layers `buffer`), the diagnostics' ntskip / ngskip, and the
    plasma's species.  field_update "accumulate" (the reference's update!,
    which never clears Exy) or "replace"; boris "reference"
    (ElectrostaticBoris) or "textbook" (ElectromagneticBoris with Ez = 0).
    field "lorenz_staggered" builds a LorenzGaugeStaggeredField instead, with
    the ImEx theta (0: Explicit(), 1: Implicit(); src/2D3V.jl uses ImEx(1));
    "lorenz" a LorenzGaugeField, run with loop() only; "lorenz_semi_implicit"
    a LorenzGaugeSemiImplicitField with fieldimex theta, sourceimex
    theta_source and the convergence test's rtol and maxiters (the
    reference's defaults sqrt(eps()) and 10).  precision "mixed"
    (electrostatic field only) runs the particles with fixed-point positions,
    FP32 hi + lo velocities, FP32 gathers and exact integer deposits; the
    field, the solve and the diagnostics stay FP64."""
    NX: int
    NY: int
    dt: float
    species: List[SpeciesConfig]
    Lx: float = 1.0
    Ly: float = 1.0
    B0: tuple = (0.0, 0.0, 0.0)
    buffer: int = 3
    ntskip: int = 1
    ngskip: int = 1
    field_update: str = "accumulate"
    boris: str = "reference"
    device: int = 0
    field: str = "electrostatic"
    theta: float = 1.0
    theta_source: float = 0.0
    rtol: float = float(np.sqrt(np.finfo(float).eps))
    maxiters: int = 10
    precision: str = "fp64"

    def to_c(self):
        c = L.pic2d3v_config()
        c.NX, c.NY = int(self.NX), int(self.NY)
        c.Lx, c.Ly, c.dt = float(self.Lx), float(self.Ly), float(self.dt)
        for k in range(3):
            c.B0[k] = float(self.B0[k])
        c.buffer, c.ntskip, c.ngskip = int(self.buffer), int(self.ntskip), int(self.ngskip)
        c.field_update = _code(FIELD_UPDATES, self.field_update)
        c.boris = _code(BORIS, self.boris)
        c.nspecies = len(self.species)
        for s, sp in enumerate(self.species[:L.MAX_SPECIES_2D3V]):
            c.species[s].P = int(sp.P)
            c.species[s].shape = _code(SHAPES, sp.shape)
            c.species[s].charge, c.species[s].mass = float(sp.charge), float(sp.mass)
            c.species[s].weight = float(sp.weight)
        c.device = int(self.device)
        c.field = _code(FIELDS, self.field)
        c.theta = float(self.theta)
        c.theta_source = float(self.theta_source)
        c.rtol = float(self.rtol)
        c.maxiters = int(self.maxiters)
        c.precision = {"fp64": L.FP64, "mixed": L.MIXED}[self.precision] if isinstance(self.precision, str) \
            else int(self.precision)
        return c


class Context2D3V:
    def __init__(self, cfg: Config2D3V):
        self._lib = L.load()
        self.cfg = cfg
        h = ctypes.c_void_p()
        c = cfg.to_c()
        rc = self._lib.pic2d3v_create(ctypes.byref(h), ctypes.byref(c))
        if rc != 0:
            raise L.PicError(rc, f"pic2d3v_create: {self._lib.pic2d3v_last_error(None).decode()}")
        self._h = h
        self.t = 0
        self._last_steps = 0

    def _check(self, rc, what):
        if rc != 0:
            raise L.PicError(rc, f"{what}: {self._lib.pic2d3v_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pic2d3v_destroy(self._h)
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

    # ---- state -------------------------------------------------------------
    def upload(self, s, xyv):
        a = np.asfortranarray(xyv, dtype=np.float64)
        if a.shape != (5, self.cfg.species[s].P):
            raise ValueError(f"xyv of species {s} must be (5, {self.cfg.species[s].P})")
        self._check(self._lib.pic2d3v_upload(self._h, int(s), L.ptr(a)), "pic2d3v_upload")

    def download(self, s):
        a = np.zeros((5, self.cfg.species[s].P), order="F")
        self._check(self._lib.pic2d3v_download(self._h, int(s), L.ptr(a)), "pic2d3v_download")
        return a

    def reset(self):
        self._check(self._lib.pic2d3v_reset(self._h), "pic2d3v_reset")
        self.t = 0

    def sort(self, s, dx, dy):
        """sort!(species, dx, dy); returns the 0-based permutation (the
        reference's species.p minus 1)."""
        perm = np.zeros(self.cfg.species[s].P, dtype=np.int64)
        self._check(self._lib.pic2d3v_sort(self._h, int(s), float(dx), float(dy), L.ptr(perm)), "pic2d3v_sort")
        return perm

    def run(self, nsteps):
        """nsteps of loop! + diagnose!; returns the rows of this call (one per t
        with t % ntskip == 0): kinetic energy, field energy, momentum (3),
        characteristic momentum (3); the Lorenz gauge adds the field momentum
        (3) before the characteristic momentum."""
        nt = self.cfg.ntskip
        t0, t1 = self.t, self.t + int(nsteps)
        n = -(-t1 // nt) - (-(-t0 // nt))
        rows = np.zeros((max(n, 0), 11 if self.cfg.field in ("lorenz_staggered", "lorenz_semi_implicit") else 8))
        rc = self._lib.pic2d3v_run(self._h, int(nsteps), L.ptr(rows))
        self.t = t1
        self._last_steps = int(nsteps)
        self._check(rc, "pic2d3v_run")
        return rows

    def fields(self, nbins=None):
        """(Exs, Eys, phis), each (nbins, NY/ngskip, NX/ngskip)."""
        if nbins is None:
            nbins = -(-self.t // self.cfg.ntskip)
        sh = (int(nbins), self.cfg.NY // self.cfg.ngskip, self.cfg.NX // self.cfg.ngskip)
        out = tuple(np.zeros(sh) for _ in range(3))
        self._check(self._lib.pic2d3v_fields(self._h, *(L.ptr(a) for a in out), int(nbins)), "pic2d3v_fields")
        return out

    def field(self):
        """The gathered field Exy: (Ex, Ey), each (NY, NX)."""
        Ex, Ey = np.zeros((self.cfg.NY, self.cfg.NX)), np.zeros((self.cfg.NY, self.cfg.NX))
        self._check(self._lib.pic2d3v_field(self._h, L.ptr(Ex), L.ptr(Ey)), "pic2d3v_field")
        return Ex, Ey

    # ---- Lorenz gauge --------------------------------------------------------
    def warmup(self):
        """warmup!(field, plasma) of a Lorenz-gauge field (nothing for
        LorenzGaugeField and the electrostatic field)."""
        self._check(self._lib.pic2d3v_warmup(self._h), "pic2d3v_warmup")

    def loop(self, nsteps):
        """nsteps of loop! without diagnose!."""
        rc = self._lib.pic2d3v_loop(self._h, int(nsteps))
        self._check(rc, "pic2d3v_loop")
        self.t += int(nsteps)
        self._last_steps = int(nsteps)

    def _slot(self, which, table):
        return table.index(which) if isinstance(which, str) else int(which)

    def _states(self):
        return L.LG_STATE_SI if self.cfg.field == "lorenz_semi_implicit" else L.LG_STATE

    def lg_get(self, which):
        """The state slot `which` (a name of LG_STATE, or of LG_STATE_SI for
        the semi-implicit field, or its index) as a complex (NY, NX) array."""
        z = np.zeros((self.cfg.NY, self.cfg.NX), dtype=np.complex128)
        self._check(self._lib.pic2d3v_lg_get(self._h, self._slot(which, self._states()), L.ptr(z)),
                    "pic2d3v_lg_get")
        return z

    def lg_state(self):
        """All slots (18, or 22 for the semi-implicit field) as a complex
        (n, NY, NX) array."""
        return np.stack([self.lg_get(k) for k in range(len(self._states()))])

    def lg_set(self, which, z):
        z = np.ascontiguousarray(np.asarray(z, dtype=np.complex128).reshape(self.cfg.NY, self.cfg.NX))
        self._check(self._lib.pic2d3v_lg_set(self._h, self._slot(which, self._states()), L.ptr(z)),
                    "pic2d3v_lg_set")

    def _grid_shape(self):
        b = self.cfg.buffer
        return (self.cfg.NY + 2 * b, self.cfg.NX + 2 * b, 4)

    def lg_grid(self):
        """The semi-implicit field's deposit grid rhoJs+: (NY+2b, NX+2b, 4),
        component fastest as rhoJs[c, i, j]."""
        out = np.zeros(self._grid_shape())
        self._check(self._lib.pic2d3v_lg_grid(self._h, L.ptr(out)), "pic2d3v_lg_grid")
        return out

    def lg_set_grid(self, grid):
        g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64).reshape(self._grid_shape()))
        self._check(self._lib.pic2d3v_lg_set_grid(self._h, L.ptr(g)), "pic2d3v_lg_set_grid")

    def lg_iterations(self, n=None):
        """The semi-implicit iteration counts of the last n steps (all steps
        of the most recent run/loop call when None)."""
        if n is None:
            n = self._last_steps
        out = np.zeros(max(int(n), 1), dtype=np.int64)
        self._check(self._lib.pic2d3v_lg_iterations(self._h, L.ptr(out), int(n)), "pic2d3v_lg_iterations")
        return out[:int(n)]

    def lg_bins(self, which=None, nbins=None):
        """The binned field `which` (a name of LG_BINS or its index), shape
        (nbins, NY/ngskip, NX/ngskip); all fourteen when which is None."""
        if nbins is None:
            nbins = -(-self.t // self.cfg.ntskip)
        if which is None:
            return np.stack([self.lg_bins(k, nbins) for k in range(len(L.LG_BINS))])
        out = np.zeros((int(nbins), self.cfg.NY // self.cfg.ngskip, self.cfg.NX // self.cfg.ngskip))
        self._check(self._lib.pic2d3v_lg_bins(self._h, self._slot(which, L.LG_BINS), L.ptr(out), int(nbins)),
                    "pic2d3v_lg_bins")
        return out

    # ---- operators ---------------------------------------------------------
    def deposit(self, shape, x, y, w):
        x, y = L.f64(x), L.f64(y)
        rho = np.zeros((self.cfg.NY, self.cfg.NX))
        self._check(self._lib.pic2d3v_deposit(self._h, _code(SHAPES, shape), L.ptr(x), L.ptr(y), x.size, float(w),
                                              L.ptr(rho)), "pic2d3v_deposit")
        return rho

    def gather(self, shape, Ex, Ey, x, y):
        Ex, Ey, x, y = L.f64(Ex), L.f64(Ey), L.f64(x), L.f64(y)
        ex, ey = np.zeros(x.size), np.zeros(x.size)
        self._check(self._lib.pic2d3v_gather(self._h, _code(SHAPES, shape), L.ptr(Ex), L.ptr(Ey), L.ptr(x), L.ptr(y),
                                             x.size, L.ptr(ex), L.ptr(ey)), "pic2d3v_gather")
        return ex, ey

    def boris(self, vx, vy, vz, ex, ey, q_m):
        """The context's Boris in place on float64 arrays vx, vy, vz."""
        for a in (vx, vy, vz):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
                raise TypeError("vx, vy, vz must be contiguous float64 arrays (updated in place)")
        ex, ey = L.f64(ex), L.f64(ey)
        self._check(self._lib.pic2d3v_boris(self._h, L.ptr(vx), L.ptr(vy), L.ptr(vz), L.ptr(ex), L.ptr(ey), vx.size,
                                            float(q_m)), "pic2d3v_boris")

    def fieldsolve(self, rho):
        rho = L.f64(rho)
        Ex, Ey, phi = (np.zeros((self.cfg.NY, self.cfg.NX)) for _ in range(3))
        self._check(self._lib.pic2d3v_fieldsolve(self._h, L.ptr(rho), L.ptr(Ex), L.ptr(Ey), L.ptr(phi)),
                    "pic2d3v_fieldsolve")
        return Ex, Ey, phi


def halton(i, base, seed=0.0):
    """halton(i, base, seed) of src/PIC2D3V.jl:29-37."""
    result, f = 0.0, 1.0
    while i > 0:
        f = f / base
        result += f * (i % base)
        i //= base
    return (result + seed) % 1.0


def species_xyv(P, vth, Lx=1.0, Ly=1.0):
    """The particle layout of Species(P, vth, density, shape; Lx, Ly)
    (src/PIC2D3V.jl:191-208): Halton positions in bases 2 and 3 with seed
    1/sqrt(2), erfinv velocities in bases 5, 7, 9, each shifted to zero mean
    and scaled to standard deviation vth/sqrt(2).  Host-side set-up; erfinv is
    scipy's."""
    from scipy.special import erfinv
    s = 1 / np.sqrt(2)
    sample = lambda b: np.array([halton(i, b, s) for i in range(P)])
    x, y = Lx * sample(2), Ly * sample(3)
    vs = []
    for b in (5, 7, 9):
        v = vth * erfinv(2 * sample(b) - 1) * vth
        v = v - v.mean()
        v = v * ((vth / np.sqrt(2)) / v.std(ddof=1))
        vs.append(v)
    return np.vstack([x, y] + vs)


def weight(n0, P, Lx, Ly):
    """calculateweight(n0, P, Lx, Ly) = n0*Lx*Ly/P (src/PIC2D3V.jl:189)."""
    return n0 * Lx * Ly / P
