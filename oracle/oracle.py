"""ctypes loader for the CPU oracle (oracle/libpicoracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker.  The product package never
imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("PICORACLE_LIB", os.path.join(_HERE, "libpicoracle.so"))

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_ip32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def build():
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    c_d, c_i32, c_i64, c_vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    sig = {
        "orc_jl_mod1f": (c_d, [c_d]),
        "orc_jl_mod1i": (c_i64, [c_i64, c_i64]),
        "orc_jl_round": (c_d, [c_d]),
        "orc_exp": (c_d, [c_d]),
        "orc_erf": (c_d, [c_d]),
        "orc_tree_sum": (c_d, [_dp, c_i64]),
        "orc_norm2": (c_d, [_dp, c_i64]),
        "orc_isapprox": (ctypes.c_int, [_dp, _dp, c_i64, c_d, c_d]),
        "orc_fxp_frac_bits": (ctypes.c_int, [c_i64, c_d]),
        "orc_fxp_from_double": (None, [c_d, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
        "orc_fxp_to_double": (c_d, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]),
        "orc_twiddles": (None, [c_i32, _dp, _dp]),
        "orc_fft": (None, [c_i32, _dp, _dp, c_i32]),
        "orc_dft_bruteforce": (None, [c_i32, _dp, _dp, _dp, _dp, c_i32]),
        "orc_fieldsolve": (None, [c_i32, _dp, _dp]),
        "orc_ngp_cell": (c_i64, [c_d, c_i32]),
        "orc_gauss_stencil": (None, [c_i32, c_i32, c_d, _ip64, _dp]),
        "orc_deposit": (None, [c_i32, c_i32, c_i64, _dp, c_vp, c_d, c_i32, _dp]),
        "orc_drift_half": (None, [c_i64, _dp, _dp, c_d]),
        "orc_kick": (None, [c_i32, c_i32, c_i64, _dp, _dp, _dp, c_d]),
        "orc_midpoint_position": (None, [c_i64, _dp, _dp, _dp, _dp, c_d]),
        "orc_push_midpoint": (None, [c_i32, c_i32, c_i64, _dp, _dp, _dp, _dp, _dp, c_d]),
        "orc_wrap": (None, [c_i64, _dp]),
        "orc_diagnostics": (None, [c_i32, c_i64, _dp, _dp, c_d, _dp]),
        "orc_quiet_start": (None, [c_i64, _dp, _dp]),
        "orc_run": (ctypes.c_int, [c_i32, c_i32, c_i32, c_i64, c_d, c_d, c_d, c_d, c_i32,
                                   _dp, _dp, _dp, c_i64, c_vp, c_vp, c_i32, c_i32]),
        "orc_area_stencil": (None, [c_i32, c_d, _ip64, _dp]),
        "orc_gather_at": (c_d, [c_i32, c_i32, _dp, c_d]),
        "orc_deposit_par": (None, [c_i32, c_i32, c_i64, _dp, c_vp, c_d, _dp, c_i32]),
        "orc_run_simpson": (ctypes.c_int, [c_i32, c_i32, c_i64, c_d, c_d, c_d, c_d, c_i32, _dp, _dp, _dp,
                                           c_i64, c_vp, c_vp, c_i32]),
        "orc_boris12": (None, [ctypes.POINTER(c_d), ctypes.POINTER(c_d), c_d, c_d, c_d, c_d]),
        "orc_deposit_1d2v": (None, [c_i32, c_i32, c_i32, c_i64, _dp, c_d, _dp, _dp]),
        "orc_run_1d2v": (ctypes.c_int, [c_i32, c_i32, c_i32, c_i64, _dp, _dp, _dp, c_d, c_d, c_d, c_d, c_i32,
                                        _dp, _dp, _dp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i32]),
        "orc2d_unimod": (c_d, [c_d, c_d]),
        "orc2d_unimodi": (c_i64, [c_i64, c_i64]),
        "orc2d_area_weights": (None, [c_d, c_i32, c_i32, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32),
                                      ctypes.POINTER(c_d), ctypes.POINTER(c_d)]),
        "orc2d_gather": (None, [c_i32, c_i32, _dp, _dp, c_d, c_d, c_i32, ctypes.POINTER(c_d),
                                ctypes.POINTER(c_d)]),
        "orc2d_boris": (None, [ctypes.POINTER(c_d), ctypes.POINTER(c_d), ctypes.POINTER(c_d), c_d, c_d, c_d,
                               c_d]),
        "orc2d_push": (None, [c_i32, c_i32, c_i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, c_d, c_d, c_i32, c_i32]),
        "orc2d_deposit": (None, [c_i32, c_i32, c_i64, _dp, _dp, c_d, c_i32, c_i32, _dp]),
        "orc2d_fft": (None, [c_i32, c_i32, _dp, _dp, c_i32]),
        "orc2d_fieldsolve": (None, [c_i32, c_i32, _dp, _dp, _dp, c_vp]),
        "orc2d_diagnostics": (None, [c_i32, c_i32, c_i64, _dp, _dp, _dp, _dp, c_d, _dp]),
        "orc2d_run": (ctypes.c_int, [c_i32, c_i32, c_i64, c_d, c_d, c_d, c_i32, c_i32, _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp, c_i64, c_vp, c_i32]),
        "orc2d_fieldsolve_k": (ctypes.c_int, [c_i32, c_i32, c_d, c_d, _dp, _dp, _dp, c_vp]),
        "orc23_unimod": (c_d, [c_d, c_d]),
        "orc23_bspline": (ctypes.c_int, [c_i32, c_d, _dp]),
        "orc23_weights": (ctypes.c_int, [c_i32, c_d, c_d, _ip64, _dp]),
        "orc23_gather": (None, [c_i32, c_i32, c_d, c_d, c_i32, _dp, _dp, c_d, c_d, ctypes.POINTER(c_d),
                                ctypes.POINTER(c_d)]),
        "orc23_boris": (None, [c_i32, _dp, c_d, c_d, c_d, c_d, ctypes.POINTER(c_d), ctypes.POINTER(c_d),
                               ctypes.POINTER(c_d)]),
        "orc23_frac_bits": (ctypes.c_int, [c_i32, _ip64, _dp]),
        "orc23_deposit": (None, [c_i32, c_i32, c_d, c_d, c_i32, c_i64, _dp, _dp, c_d, _dp]),
        "orc23_fieldsolve": (ctypes.c_int, [c_i32, c_i32, c_d, c_d, _dp, _dp, _dp, c_vp]),
        "orc23_order_key": (ctypes.c_uint64, [c_d]),
        "orc23_ceil_index": (c_i64, [c_d]),
        "orc23_sort": (None, [c_i64, _dp, _dp, _dp, _dp, _dp, c_d, c_d, _ip64]),
        "orc23_diagnostics": (None, [c_i32, c_i32, c_i32, _dp, _dp, c_i32, _ip64, _dp, _dp, _dp, _dp]),
        "orc_em_irn": (c_i64, [c_d, c_i32]),
        "orc_em_stencil": (None, [c_i32, c_d, _ip64, _dp]),
        "orc_em_frac_bits": (ctypes.c_int, [c_i64, c_d, c_d]),
        "orc_em_deposit": (None, [c_i32, c_i64, _dp, c_vp, c_d, _dp]),
        "orc_em_fields": (None, [c_i32, c_d, c_d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_em_push": (None, [c_i32, c_d, _dp, _dp, _dp, _dp, _dp, c_d, _dp, _dp]),
        "orc_em_sort": (None, [c_i64, _dp, _dp, _dp, _dp, _ip64]),
        "orc_em_run": (ctypes.c_int, [c_i32, c_i64, c_d, c_d, c_d, c_d, c_i32, c_i32, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, c_i64, c_vp, c_vp]),
        "orc_emx_gauss_faraday": (None, [c_i32, c_d, _dp, _dp, _dp, _dp]),
        "orc_emx_ampere": (None, [c_i32, c_d, c_d, _dp, _dp, _dp, _dp]),
        "orc_emx_push": (None, [c_i32, c_d, _dp, _dp, c_d, _dp]),
        "orc_emx_run": (ctypes.c_int, [c_i32, c_i64, c_d, c_d, c_d, c_i32, _dp, _dp, _dp, _dp, _dp, _dp, c_i64,
                                       c_vp]),
        "orclg_wavenumbers": (None, [c_i32, c_i32, c_d, c_d, _dp, _dp, _dp]),
        "orclg_boris": (None, [c_d, c_d, _dp, _dp, _dp]),
        "orclg_gather": (None, [c_i32, c_i32, c_d, c_d, c_i32, _dp, c_d, c_d, _dp]),
        "orclg_ifft2": (None, [c_i32, c_i32, _dp]),
        "orclg_solve": (None, [c_i32, c_i32, c_d, c_d, c_d, c_d, _dp, _dp]),
        "orclg_update": (None, [c_i32, c_i32, _dp, _dp, _dp]),
        "orclg_frac_bits": (ctypes.c_int, [c_i32, _ip64, _dp, c_vp]),
        "orclg_warmup": (None, [c_i32, c_i32, c_d, c_d, c_d, c_i32, _ip32, _ip64, _dp, _dp, _dp, _dp]),
        "orclg_row": (None, [c_i32, c_i32, c_i32, _dp, _dp, c_i32, _ip64, _dp, _dp, _dp, _dp]),
        "orclg_fft2": (None, [c_i32, c_i32, _dp]),
        "orclg_solve_centred": (None, [c_i32, c_i32, c_d, c_d, c_d, c_d, _dp]),
        "orclg_run_centred": (ctypes.c_int, [c_i32, c_i32, c_d, c_d, c_d, _dp, c_d, c_i32, _ip32, _ip64, _dp, _dp,
                                             _dp, _dp, _dp, _dp, c_i64]),
        "orclg_run": (ctypes.c_int, [c_i32, c_i32, c_d, c_d, c_d, _dp, c_i32, c_d, c_i32, c_i32, c_i32, _ip32,
                                     _ip64, _dp, _dp, _dp, _dp, _dp, c_i64, c_i64, c_i32, c_vp, c_vp, c_i64]),
        "orclg_si_deposit": (ctypes.c_int, [c_i32, c_i32, c_i32, c_d, c_d, c_i32, _ip32, _ip64, _dp, _dp, _dp,
                                            c_i32, _dp, _dp]),
        "orclg_si_warmup": (ctypes.c_int, [c_i32, c_i32, c_d, c_d, c_d, c_i32, c_i32, _ip32, _ip64, _dp, _dp, _dp,
                                           _dp, _dp]),
        "orclg_si_solve": (None, [c_i32, c_i32, c_d, c_d, c_d, c_d, c_d, _dp, _dp, _dp]),
        "orclg_si_run": (ctypes.c_int, [c_i32, c_i32, c_d, c_d, c_d, _dp, c_i32, c_d, c_d, c_d, c_i64, c_i32, c_i32,
                                        c_i32, _ip32, _ip64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, c_i64, c_i64, c_i32,
                                        c_vp, c_vp, c_i64, _ip64]),
        "orc23_run": (ctypes.c_int, [c_i32, c_i32, c_d, c_d, c_d, _dp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32,
                                     _ip32, _ip64, _dp, _dp, _dp, _dp, _dp, _dp, c_i64, c_i64, c_vp, c_vp, c_vp,
                                     c_vp, c_i64, c_i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def jl_mod1f(x): return lib().orc_jl_mod1f(float(x))
def jl_mod1i(i, n): return lib().orc_jl_mod1i(int(i), int(n))
def jl_round(x): return lib().orc_jl_round(float(x))
def exp(x): return lib().orc_exp(float(x))
def erf(x): return lib().orc_erf(float(x))


def erf_array(xs):
    L = lib()
    return np.array([L.orc_erf(float(v)) for v in np.asarray(xs).ravel()]).reshape(np.shape(xs))


def exp_array(xs):
    L = lib()
    return np.array([L.orc_exp(float(v)) for v in np.asarray(xs).ravel()]).reshape(np.shape(xs))


def tree_sum(x):
    x = _f64(x)
    return lib().orc_tree_sum(x, x.size)


def norm2(x):
    x = _f64(x)
    return lib().orc_norm2(x, x.size)


def isapprox(F, E, rtol, atol=0.0):
    F, E = _f64(F), _f64(E)
    return bool(lib().orc_isapprox(F, E, F.size, float(rtol), float(atol)))


def fxp_frac_bits(P, w):
    return lib().orc_fxp_frac_bits(int(P), float(w))


def fxp_from_double(a, q):
    lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_fxp_from_double(float(a), int(q), ctypes.byref(lo), ctypes.byref(hi))
    return (hi.value << 64) | lo.value


def fxp_to_double(S, q):
    return lib().orc_fxp_to_double(S & ((1 << 64) - 1), S >> 64, int(q))


def twiddles(N):
    re, im = np.zeros(N // 2), np.zeros(N // 2)
    lib().orc_twiddles(N, re, im)
    return re + 1j * im


def fft(z, inverse=False):
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real).copy(), _f64(z.imag).copy()
    lib().orc_fft(z.size, re, im, 1 if inverse else 0)
    return re + 1j * im


def dft_bruteforce(z, inverse=False):
    z = np.asarray(z, dtype=np.complex128)
    N = z.size
    orr, oi = np.zeros(N), np.zeros(N)
    lib().orc_dft_bruteforce(N, _f64(z.real).copy(), _f64(z.imag).copy(), orr, oi, 1 if inverse else 0)
    return orr + 1j * oi


def fieldsolve(rho):
    rho = _f64(rho)
    E = np.zeros_like(rho)
    lib().orc_fieldsolve(rho.size, rho, E)
    return E


def ngp_cell(x, N):
    return lib().orc_ngp_cell(float(x), int(N))


def gauss_stencil(N, R, c):
    cells = np.zeros(2 * R + 1, dtype=np.int64)
    wt = np.zeros(2 * R + 1)
    lib().orc_gauss_stencil(int(N), int(R), float(c), cells, wt)
    return cells, wt


def deposit(N, shape, x, y=None, w=1.0, mode=0):
    x = _f64(x)
    y = None if y is None else _f64(y)
    rho = np.zeros(N)
    lib().orc_deposit(int(N), int(shape), x.size, x, _ptr(y), float(w), int(mode), rho)
    return rho


def drift_half(x, v, dt):
    x = _f64(x).copy()
    lib().orc_drift_half(x.size, x, _f64(v), float(dt))
    return x


def kick(N, shape, v, x, E, dt):
    v = _f64(v).copy()
    lib().orc_kick(int(N), int(shape), v.size, v, _f64(x), _f64(E), float(dt))
    return v


def midpoint_position(X, V, v, dt):
    X = _f64(X)
    x = np.zeros_like(X)
    lib().orc_midpoint_position(X.size, X, _f64(V), _f64(v), x, float(dt))
    return x


def push_midpoint(N, R, X, V, x, E, dt):
    X = _f64(X)
    v = np.zeros_like(X)
    lib().orc_push_midpoint(int(N), int(R), X.size, X, _f64(V), _f64(x), v, _f64(E), float(dt))
    return v


def wrap(x):
    x = _f64(x).copy()
    lib().orc_wrap(x.size, x)
    return x


def diagnostics(E, v, W):
    E, v = _f64(E), _f64(v)
    out = np.zeros(4)
    lib().orc_diagnostics(E.size, v.size, E, v, float(W), out)
    return out


def quiet_start(P):
    x, v = np.zeros(P), np.zeros(P)
    lib().orc_quiet_start(int(P), x, v)
    return x, v


def run(N, shape, scheme, W, w, dt, x, v, nsteps, rtol=1e-8, max_sweeps=10, E=None,
        deposit_mode=0, nthreads=1):
    """Run nsteps of the restated script loop.  Returns (x, v, E, diag, sweeps)."""
    x = _f64(x).copy()
    v = _f64(v).copy()
    E = np.zeros(N) if E is None else _f64(E).copy()
    diag = np.zeros((nsteps, 4))
    sweeps = np.zeros(nsteps, dtype=np.int32)
    rc = lib().orc_run(int(N), int(shape), int(scheme), x.size, float(W), float(w), float(dt),
                       float(rtol), int(max_sweeps), x, v, E, int(nsteps),
                       _ptr(diag), _ptr(sweeps), int(deposit_mode), int(nthreads))
    if rc != 0:
        raise ValueError(f"orc_run rejected its arguments (rc={rc})")
    return x, v, E, diag, sweeps


# ---- Simpson-1/3 fixed point (src/*Simpson13.jl) ----------------------------
SHAPE_AREA = 1


def area_stencil(N, c):
    """d(y) of src/AreaFixedPointQuietSimpson13.jl:5: ((cell, weight), (cell, weight)), cells 0-based."""
    cells = np.zeros(2, dtype=np.int64)
    wt = np.zeros(2)
    lib().orc_area_stencil(int(N), float(c), cells, wt)
    return (int(cells[0]), float(wt[0])), (int(cells[1]), float(wt[1]))


def gather_at(N, shape, E, c):
    """sum(k -> E[k[1]]*k[2], d(c)) as a left fold; shape 1 (area), 6 or 7 (Gaussian)."""
    return lib().orc_gather_at(int(N), int(shape), _f64(E), float(c))


def run_simpson(N, shape, W, w, dt, x, v, nsteps, rtol, max_sweeps=10, E3=None, nthreads=1):
    """nsteps of the Simpson-1/3 loop.  E3 is the (3, N) field matrix (rows
    E[1,:], E[2,:], E[3,:]).  Returns (x, v, E3, diag, sweeps)."""
    x = _f64(x).copy()
    v = _f64(v).copy()
    Ecm = np.zeros(3 * N) if E3 is None else np.ascontiguousarray(_f64(E3).T).ravel().copy()
    diag = np.zeros((nsteps, 4))
    sweeps = np.zeros(nsteps, dtype=np.int32)
    rc = lib().orc_run_simpson(int(N), int(shape), x.size, float(W), float(w), float(dt), float(rtol),
                               int(max_sweeps), x, v, Ecm, int(nsteps), _ptr(diag), _ptr(sweeps), int(nthreads))
    if rc != 0:
        raise ValueError(f"orc_run_simpson rejected its arguments (rc={rc})")
    return x, v, Ecm.reshape(N, 3).T.copy(), diag, sweeps


# ---- magnetised 1D2V (src/NGP1D2V.jl, src/NGP1D2V2S.jl) ----------------------

def boris12(vx, vy, E, B, dt, q_m=1.0):
    a, b = ctypes.c_double(vx), ctypes.c_double(vy)
    lib().orc_boris12(ctypes.byref(a), ctypes.byref(b), float(E), float(B), float(dt), float(q_m))
    return a.value, b.value


def deposit_1d2v(N, R, q, w, xs):
    """rho() over the species position arrays xs (charges q), exact."""
    ns = len(xs)
    x = np.concatenate([_f64(a) for a in xs])
    rho = np.zeros(N)
    lib().orc_deposit_1d2v(int(N), int(R), ns, x.size // ns, _f64(np.asarray(q, dtype=np.float64)), float(w),
                           x, rho)
    return rho


def run_1d2v(N, R, q, q_m, mass, w, n0, B0, dt, bin, xs, vxs, vys, nsteps, t0=0, nbins=None, Es=None, D=None,
             nthreads=1):
    """nsteps of the 1D2V loop for the species arrays xs, vxs, vys (lists of
    equal length P).  Returns (xs, vxs, vys, Es (nbins, N), D (nbins, 5)), Es
    and D before the scripts' division by bin."""
    ns = len(xs)
    P = _f64(xs[0]).size
    x = np.concatenate([_f64(a) for a in xs])
    vx = np.concatenate([_f64(a) for a in vxs])
    vy = np.concatenate([_f64(a) for a in vys])
    if nbins is None:
        nbins = (t0 + nsteps + bin - 1) // bin
    Es = np.zeros((nbins, N)) if Es is None else np.ascontiguousarray(_f64(Es)).copy()
    D = np.zeros((nbins, 5)) if D is None else np.ascontiguousarray(_f64(D)).copy()
    arr = lambda a: _f64(np.asarray(a, dtype=np.float64))
    rc = lib().orc_run_1d2v(int(N), int(R), ns, P, arr(q), arr(q_m), arr(mass), float(w), float(n0), float(B0),
                            float(dt), int(bin), x, vx, vy, int(t0), int(nsteps), _ptr(Es), int(nbins), _ptr(D),
                            int(nthreads))
    if rc != 0:
        raise ValueError(f"orc_run_1d2v rejected its arguments (rc={rc})")
    split = lambda a: [a[s * P:(s + 1) * P].copy() for s in range(ns)]
    return split(x), split(vx), split(vy), Es, D


# ---- electrostatic 2D3V (src/Electrostatic2D3V.jl) -------------------------
# Grids are (NY, NX) numpy arrays in C order, i.e. Julia's column-major NX x NY
# with x fastest; flat index i + NX*j.

def area_weights(z, NZ, textbook=0):
    c0, c1, w0, w1 = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
    lib().orc2d_area_weights(float(z), int(NZ), int(textbook), ctypes.byref(c0), ctypes.byref(c1),
                             ctypes.byref(w0), ctypes.byref(w1))
    return (c0.value, w0.value), (c1.value, w1.value)


def gather2d(Ex, Ey, x, y, textbook=0):
    Ex, Ey = _f64(Ex), _f64(Ey)
    NY, NX = Ex.shape
    ex, ey = ctypes.c_double(), ctypes.c_double()
    lib().orc2d_gather(NX, NY, Ex.ravel(), Ey.ravel(), float(x), float(y), int(textbook),
                       ctypes.byref(ex), ctypes.byref(ey))
    return ex.value, ey.value


def boris(vx, vy, vz, Ex, Ey, dt, B0):
    a, b, c = ctypes.c_double(vx), ctypes.c_double(vy), ctypes.c_double(vz)
    lib().orc2d_boris(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), float(Ex), float(Ey), float(dt),
                      float(B0))
    return a.value, b.value, c.value


def push2d(x, y, vx, vy, vz, Ex, Ey, dt, B0, textbook=0, nthreads=1):
    """Gather, Boris and move every particle; returns new (x, y, vx, vy, vz)."""
    x, y, vx, vy, vz = (_f64(a).copy() for a in (x, y, vx, vy, vz))
    Ex, Ey = _f64(Ex), _f64(Ey)
    NY, NX = Ex.shape
    lib().orc2d_push(NX, NY, x.size, x, y, vx, vy, vz, Ex.ravel(), Ey.ravel(), float(dt), float(B0),
                     int(textbook), int(nthreads))
    return x, y, vx, vy, vz


def deposit2d(NX, NY, x, y, w, textbook=0, nthreads=1):
    x, y = _f64(x), _f64(y)
    rho = np.zeros((NY, NX))
    lib().orc2d_deposit(int(NX), int(NY), x.size, x, y, float(w), int(textbook), int(nthreads), rho.ravel())
    return rho


def fft2d(re, im, inverse=0):
    re, im = _f64(re).copy(), _f64(im).copy()
    NY, NX = re.shape
    lib().orc2d_fft(NX, NY, re.ravel(), im.ravel(), int(inverse))
    return re, im


def fieldsolve2d(rho, want_phi=False):
    rho = _f64(rho)
    NY, NX = rho.shape
    Ex, Ey = np.zeros((NY, NX)), np.zeros((NY, NX))
    phi = np.zeros((NY, NX)) if want_phi else None
    lib().orc2d_fieldsolve(NX, NY, rho.ravel(), Ex.ravel(), Ey.ravel(), _ptr(phi))
    return (Ex, Ey, phi) if want_phi else (Ex, Ey)


def diagnostics2d(Ex, Ey, vx, vy, w):
    Ex, Ey, vx, vy = _f64(Ex), _f64(Ey), _f64(vx), _f64(vy)
    NY, NX = Ex.shape
    out = np.zeros(5)
    lib().orc2d_diagnostics(NX, NY, vx.size, Ex.ravel(), Ey.ravel(), vx, vy, float(w), out)
    return out


def run2d(NX, NY, dt, B0, w, x, y, vx, vy, vz, nsteps, ns=2, textbook=0, Ex=None, Ey=None, nthreads=1):
    """nsteps of the restated 2D3V loop.  Returns (x, y, vx, vy, vz, Ex, Ey, K)
    with K of shape (nsteps // ns, 5)."""
    x, y, vx, vy, vz = (_f64(a).copy() for a in (x, y, vx, vy, vz))
    Ex = np.zeros((NY, NX)) if Ex is None else _f64(Ex).copy()
    Ey = np.zeros((NY, NX)) if Ey is None else _f64(Ey).copy()
    K = np.zeros((nsteps // ns, 5))
    rc = lib().orc2d_run(int(NX), int(NY), x.size, float(dt), float(B0), float(w), int(textbook), int(ns),
                         x, y, vx, vy, vz, Ex.ravel(), Ey.ravel(), int(nsteps), _ptr(K), int(nthreads))
    if rc != 0:
        raise ValueError(f"orc2d_run rejected its arguments (rc={rc})")
    return x, y, vx, vy, vz, Ex, Ey, K


# ---- structured 2D3V module, electrostatic path (src/PIC2D3V.jl) ------------
# Grids are (NY, NX) numpy arrays in C order (Julia's NX x NY, x fastest).
# Particles of one species are a (5, P) array of rows x, y, vx, vy, vz (the
# transpose of the reference's 5 x P xyv matrix as numpy sees it).

SHAPES23 = {"ngp": 0, "area": 1, **{f"bspline{n}": 2 + n for n in range(6)}}


def _shape23(shape):
    return SHAPES23[shape] if isinstance(shape, str) else int(shape)


def unimod23(x, n):
    return lib().orc23_unimod(float(x), float(n))


def bspline(n, x):
    w = np.zeros(6)
    m = lib().orc23_bspline(int(n), float(x), w)
    return tuple(w[:m])


def weights23(shape, z, NZ_Lz):
    """(Julia cell indices, weights) of depositindicesfractions."""
    c, w = np.zeros(6, dtype=np.int64), np.zeros(6)
    m = lib().orc23_weights(_shape23(shape), float(z), float(NZ_Lz), c, w)
    return list(c[:m]), list(w[:m])


def gather23(Ex, Ey, x, y, shape, Lx=1.0, Ly=1.0):
    Ex, Ey = _f64(Ex), _f64(Ey)
    NY, NX = Ex.shape
    ex, ey = ctypes.c_double(), ctypes.c_double()
    lib().orc23_gather(NX, NY, NX / Lx, NY / Ly, _shape23(shape), Ex.ravel(), Ey.ravel(), float(x), float(y),
                       ctypes.byref(ex), ctypes.byref(ey))
    return ex.value, ey.value


def boris23(vx, vy, vz, Ex, Ey, B, dt, q_m, variant=0):
    a, b, c = ctypes.c_double(vx), ctypes.c_double(vy), ctypes.c_double(vz)
    lib().orc23_boris(int(variant), _f64(B), float(dt), float(q_m), float(Ex), float(Ey), ctypes.byref(a),
                      ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def frac_bits23(P, w):
    P, w = np.ascontiguousarray(P, dtype=np.int64), _f64(w)
    return lib().orc23_frac_bits(P.size, P, w)


def deposit23(NX, NY, x, y, w, shape, Lx=1.0, Ly=1.0):
    x, y = _f64(x), _f64(y)
    rho = np.zeros((NY, NX))
    lib().orc23_deposit(int(NX), int(NY), NX / Lx, NY / Ly, _shape23(shape), x.size, x, y, float(w), rho.ravel())
    return rho


def fieldsolve23(rho, Lx=1.0, Ly=1.0):
    """(Ex, Ey, phi, nonfinite) of the electrostatic solve in an Lx x Ly box."""
    rho = _f64(rho)
    NY, NX = rho.shape
    Ex, Ey, phi = np.zeros((NY, NX)), np.zeros((NY, NX)), np.zeros((NY, NX))
    bad = lib().orc23_fieldsolve(NX, NY, float(Lx), float(Ly), rho.ravel(), Ex.ravel(), Ey.ravel(), _ptr(phi))
    return Ex, Ey, phi, bool(bad)


def order_key(v):
    return lib().orc23_order_key(float(v))


def sort23(xyv, dx, dy):
    """sort!(s, dx, dy): returns (sorted (5, P) copy, 0-based permutation)."""
    a = _f64(xyv).copy()
    P = a.shape[1]
    perm = np.zeros(P, dtype=np.int64)
    rows = [np.ascontiguousarray(a[k]) for k in range(5)]
    lib().orc23_sort(P, *rows, float(dx), float(dy), perm)
    return np.vstack(rows), perm


def diagnostics23(Eacc_x, Eacc_y, buffer, species):
    """Row of 8 (kinetic, field, momentum x3, characteristic momentum x3);
    species: list of dicts with mass, weight, xyv ((5, P) array)."""
    Ex, Ey = _f64(Eacc_x), _f64(Eacc_y)
    NY, NX = Ex.shape
    P = np.array([s["xyv"].shape[1] for s in species], dtype=np.int64)
    state = np.concatenate([_f64(s["xyv"]).ravel() for s in species])
    row = np.zeros(8)
    lib().orc23_diagnostics(NX, NY, int(buffer), Ex.ravel(), Ey.ravel(), len(species), P,
                            _f64([s["mass"] for s in species]), _f64([s["weight"] for s in species]), state, row)
    return row


def run23(NX, NY, Lx, Ly, dt, B, species, nsteps, t0=0, buffer=3, ntskip=1, ngskip=1, accumulate=1, boris=0,
          Eacc=None, nbins=None, bins=None, nthreads=1):
    """nsteps of loop! + diagnose! of the electrostatic PIC2D3V module.

    species: list of dicts with shape, charge, mass, weight and xyv ((5, P)).
    Eacc: (Ex, Ey) gathered field carried between calls (zeros at the start).
    bins: (Exs, Eys, phis) arrays of shape (nbins, NY/ngskip, NX/ngskip),
    accumulated in place (zeros when None).
    Returns (species with new xyv, Eacc, rows (n, 8), bins, nonfinite)."""
    P = np.array([s["xyv"].shape[1] for s in species], dtype=np.int64)
    state = np.concatenate([_f64(s["xyv"]).ravel() for s in species])
    Ex = np.zeros((NY, NX)) if Eacc is None else _f64(Eacc[0]).copy()
    Ey = np.zeros((NY, NX)) if Eacc is None else _f64(Eacc[1]).copy()
    t1 = t0 + nsteps
    nrows = -(-t1 // ntskip) - (-(-t0 // ntskip))
    rows = np.zeros((max(nrows, 0), 8))
    if nbins is None:
        nbins = -(-t1 // ntskip)
    if bins is None:
        bins = tuple(np.zeros((nbins, NY // ngskip, NX // ngskip)) for _ in range(3))
    bins = tuple(_f64(b).copy() for b in bins)
    bad = lib().orc23_run(int(NX), int(NY), float(Lx), float(Ly), float(dt), _f64(B), int(buffer), int(ntskip),
                          int(ngskip), int(accumulate), int(boris), len(species),
                          np.array([_shape23(s["shape"]) for s in species], dtype=np.int32), P,
                          _f64([s["charge"] for s in species]), _f64([s["mass"] for s in species]),
                          _f64([s["weight"] for s in species]), state, Ex.ravel(), Ey.ravel(), int(t0),
                          int(nsteps), _ptr(rows), _ptr(bins[0]), _ptr(bins[1]), _ptr(bins[2]), int(nbins),
                          int(nthreads))
    out, o = [], 0
    for s, n in zip(species, P):
        d = dict(s)
        d["xyv"] = state[o:o + 5 * n].reshape(5, n).copy()
        out.append(d)
        o += 5 * n
    return out, (Ex, Ey), rows, bins, bool(bad)


# ---- Lorenz-gauge electromagnetic 2D3V (src/PIC2D3V.jl) -------------------
# Field state: complex128 array (18, NY, NX) in the order of LG_FIELDS.
LG_FIELDS = ["phi0", "phim", "Ax0", "Ay0", "Az0", "Axm", "Aym", "Azm", "rho0", "Jx0", "Jy0", "Jz0",
             "Ex", "Ey", "Ez", "Bx", "By", "Bz"]
LG_BINS = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Ax", "Ay", "Az", "phi", "rho", "Jx", "Jy", "Jz"]


def lg_wavenumbers(NX, NY, Lx=1.0, Ly=1.0):
    kx, ky, k2 = np.zeros(NX), np.zeros(NY), np.zeros((NY, NX))
    lib().orclg_wavenumbers(int(NX), int(NY), float(Lx), float(Ly), kx, ky, k2.ravel())
    return kx, ky, k2


def lg_boris(dt, q_m, E, B, v):
    v = _f64(v).copy()
    lib().orclg_boris(float(dt), float(q_m), _f64(E), _f64(B), v)
    return v


def lg_gather(EB, x, y, shape, Lx=1.0, Ly=1.0):
    """EB: (6, NY, NX) real: Ex, Ey, Ez, Bx, By, Bz."""
    EB = _f64(EB)
    _, NY, NX = EB.shape
    out = np.zeros(6)
    lib().orclg_gather(NX, NY, NX / Lx, NY / Ly, _shape23(shape), EB.ravel(), float(x), float(y), out)
    return out


def _c128(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def _cp(a):
    return a.view(np.float64).ravel()


def lg_ifft2(z):
    z = _c128(z).copy()
    NY, NX = z.shape
    lib().orclg_ifft2(NX, NY, _cp(z))
    return z


def lg_solve(F, dt, theta, Lx=1.0, Ly=1.0):
    """Steps 1-2 of loop!: returns (F with Ex..Bz replaced, plus (4, NY, NX))."""
    F = _c128(F).copy()
    _, NY, NX = F.shape
    plus = np.zeros((4, NY, NX), dtype=np.complex128)
    lib().orclg_solve(NX, NY, float(Lx), float(Ly), float(dt), float(theta), _cp(F), _cp(plus))
    return F, plus


def lg_update(F, B0):
    F = _c128(F)
    _, NY, NX = F.shape
    EB = np.zeros((6, NY, NX))
    lib().orclg_update(NX, NY, _cp(F), _f64(B0), EB.ravel())
    return EB


def lg_frac_bits(P, w, M=None):
    P = np.asarray(P, dtype=np.int64)
    return lib().orclg_frac_bits(P.size, P, _f64(w), _ptr(None if M is None else _f64(M)))


def _lg_species(species):
    P = np.array([s["xyv"].shape[1] for s in species], dtype=np.int64)
    state = np.concatenate([_f64(s["xyv"]).ravel() for s in species])
    shape = np.array([_shape23(s["shape"]) for s in species], dtype=np.int32)
    return P, state, shape


def _lg_unpack(species, P, state):
    out, o = [], 0
    for s, n in zip(species, P):
        d = dict(s)
        d["xyv"] = state[o:o + 5 * n].reshape(5, n).copy()
        out.append(d)
        o += 5 * n
    return out


def lg_warmup(NX, NY, Lx, Ly, dt, species, F=None):
    """warmup!(::LorenzGaugeStaggeredField): returns (species, F)."""
    P, state, shape = _lg_species(species)
    F = np.zeros((18, NY, NX), dtype=np.complex128) if F is None else _c128(F).copy()
    lib().orclg_warmup(int(NX), int(NY), float(Lx), float(Ly), float(dt), len(species), shape, P,
                       _f64([s["charge"] for s in species]), _f64([s["weight"] for s in species]), state, _cp(F))
    return _lg_unpack(species, P, state), F


def lg_row(F, EB, species, buffer):
    P, state, _ = _lg_species(species)
    _, NY, NX = EB.shape
    row = np.zeros(11)
    lib().orclg_row(NX, NY, int(buffer), _cp(_c128(F)), _f64(EB).ravel(), len(species), P,
                    _f64([s["mass"] for s in species]), _f64([s["weight"] for s in species]), state, row)
    return row


def lg_run(NX, NY, Lx, Ly, dt, B0, species, F, nsteps, theta=1.0, t0=0, buffer=10, ntskip=1, ngskip=1,
           diagnose=True, nbins=None, bins=None):
    """nsteps of loop!(plasma, ::LorenzGaugeStaggeredField) and (diagnose)
    diagnose!(::LorenzGaugeDiagnostics).  F: (18, NY, NX) complex field
    state.  bins: (14, nbins, NY/ngskip, NX/ngskip), accumulated in place
    (zeros when None).  Returns (species, F, rows (n, 11), bins)."""
    P, state, shape = _lg_species(species)
    F = _c128(F).copy()
    t1 = t0 + nsteps
    nrows = (-(-t1 // ntskip) - (-(-t0 // ntskip))) if diagnose else 0
    rows = np.zeros((max(nrows, 0), 11))
    if nbins is None:
        nbins = -(-t1 // ntskip)
    if bins is None:
        bins = np.zeros((14, nbins, NY // ngskip, NX // ngskip))
    bins = _f64(bins).copy()
    rc = lib().orclg_run(int(NX), int(NY), float(Lx), float(Ly), float(dt), _f64(B0), int(buffer), float(theta),
                         int(ntskip), int(ngskip), len(species), shape, P, _f64([s["charge"] for s in species]),
                         _f64([s["mass"] for s in species]), _f64([s["weight"] for s in species]), state, _cp(F),
                         int(t0), int(nsteps), int(bool(diagnose)), _ptr(rows), _ptr(bins), int(nbins))
    if rc != 0:
        raise ValueError(f"orclg_run rejected its arguments (rc={rc})")
    return _lg_unpack(species, P, state), F, rows, bins


def lg_fft2(z):
    z = _c128(z).copy()
    NY, NX = z.shape
    lib().orclg_fft2(NX, NY, _cp(z))
    return z


def lg_solve_centred(F, dt, theta, Lx=1.0, Ly=1.0):
    """Steps 3-4 of LorenzGaugeField's loop! on the state (sources already in
    the spectrum).  Returns the new state."""
    F = _c128(F).copy()
    _, NY, NX = F.shape
    lib().orclg_solve_centred(NX, NY, float(Lx), float(Ly), float(dt), float(theta), _cp(F))
    return F


def lg_run_centred(NX, NY, Lx, Ly, dt, B0, species, F, EB, nsteps, theta=0.0):
    """nsteps of loop!(plasma, ::LorenzGaugeField).  F: (18, NY, NX) complex
    state (PHI0 = phi+, PHIM = phi0, ...); EB: (6, NY, NX) the gathered
    fields (zeros for a fresh field).  Returns (species, F, EB)."""
    P, state, shape = _lg_species(species)
    F = _c128(F).copy()
    EB = _f64(EB).copy()
    rc = lib().orclg_run_centred(int(NX), int(NY), float(Lx), float(Ly), float(dt), _f64(B0), float(theta),
                                 len(species), shape, P, _f64([s["charge"] for s in species]),
                                 _f64([s["mass"] for s in species]), _f64([s["weight"] for s in species]), state,
                                 _cp(F), EB.ravel(), int(nsteps))
    if rc != 0:
        raise ValueError(f"orclg_run_centred rejected its arguments (rc={rc})")
    return _lg_unpack(species, P, state), F, EB


# LorenzGaugeSemiImplicitField: 22 slots (LG_FIELDS + rho-, J-), the buffered
# deposit grid (NY + 2b, NX + 2b, 4) (component fastest, as rhoJs[c, i, j])
# and EB (6, NY, NX).
LG_SI_FIELDS = LG_FIELDS + ["rhom", "Jxm", "Jym", "Jzm"]


def lg_si_deposit(NX, NY, Lx, Ly, species, buffer, currents_only=False):
    """The exact deposits of the species at their positions into the
    buffered grid and the periodic cells: returns (rc, grid (NY+2b, NX+2b, 4),
    per (4, NY, NX) complex)."""
    P, state, shape = _lg_species(species)
    b = int(buffer)
    grid = np.zeros((NY + 2 * b, NX + 2 * b, 4))
    per = np.zeros((4, NY, NX), dtype=np.complex128)
    rc = lib().orclg_si_deposit(int(NX), int(NY), b, float(Lx), float(Ly), len(species), shape, P,
                                _f64([s["charge"] for s in species]), _f64([s["weight"] for s in species]), state,
                                int(bool(currents_only)), grid.ravel(), _cp(per))
    return rc, grid, per


def lg_si_warmup(NX, NY, Lx, Ly, dt, species, buffer):
    """warmup!(::LorenzGaugeSemiImplicitField): returns (species, F (22, NY, NX), grid)."""
    P, state, shape = _lg_species(species)
    b = int(buffer)
    F = np.zeros((22, NY, NX), dtype=np.complex128)
    grid = np.zeros((NY + 2 * b, NX + 2 * b, 4))
    rc = lib().orclg_si_warmup(int(NX), int(NY), float(Lx), float(Ly), float(dt), b, len(species), shape, P,
                               _f64([s["charge"] for s in species]), _f64([s["weight"] for s in species]), state,
                               _cp(F), grid.ravel())
    if rc != 0:
        raise ValueError(f"orclg_si_warmup failed (rc={rc})")
    return _lg_unpack(species, P, state), F, grid


def lg_si_solve(F, src, dt, theta_f, theta_s, Lx=1.0, Ly=1.0):
    """Steps 2-4 of an iteration: returns (F with real-space E, B; src in the
    spectrum with [1,1] = 0; plus (4, NY, NX))."""
    F = _c128(F).copy()
    src = _c128(src).copy()
    _, NY, NX = F.shape
    plus = np.zeros((4, NY, NX), dtype=np.complex128)
    lib().orclg_si_solve(NX, NY, float(Lx), float(Ly), float(dt), float(theta_f), float(theta_s), _cp(F), _cp(src),
                         _cp(plus))
    return F, src, plus


def lg_si_run(NX, NY, Lx, Ly, dt, B0, species, F, grid, EB, nsteps, theta=1.0, theta_source=0.0,
              rtol=float(np.sqrt(np.finfo(float).eps)), maxiters=10, t0=0, buffer=10, ntskip=1, ngskip=1,
              diagnose=True, nbins=None, bins=None):
    """nsteps of loop!(plasma, ::LorenzGaugeSemiImplicitField) and (diagnose)
    diagnose!.  Returns (species, F, grid, EB, rows, bins, iterations)."""
    P, state, shape = _lg_species(species)
    F = _c128(F).copy()
    grid = _f64(grid).copy()
    EB = _f64(EB).copy()
    t1 = t0 + nsteps
    nrows = (-(-t1 // ntskip) - (-(-t0 // ntskip))) if diagnose else 0
    rows = np.zeros((max(nrows, 0), 11))
    if nbins is None:
        nbins = -(-t1 // ntskip)
    if bins is None:
        bins = np.zeros((14, nbins, NY // ngskip, NX // ngskip))
    bins = _f64(bins).copy()
    iters = np.zeros(max(int(nsteps), 1), dtype=np.int64)
    rc = lib().orclg_si_run(int(NX), int(NY), float(Lx), float(Ly), float(dt), _f64(B0), int(buffer), float(theta),
                            float(theta_source), float(rtol), int(maxiters), int(ntskip), int(ngskip), len(species),
                            shape, P, _f64([s["charge"] for s in species]), _f64([s["mass"] for s in species]),
                            _f64([s["weight"] for s in species]), state, _cp(F), grid.ravel(), EB.ravel(), int(t0),
                            int(nsteps), int(bool(diagnose)), _ptr(rows), _ptr(bins), int(nbins), iters)
    if rc != 0:
        raise ValueError(f"orclg_si_run failed (rc={rc})")
    return _lg_unpack(species, P, state), F, grid, EB, rows, bins, iters[:nsteps]


# ---- 1D electromagnetic fixed point (src/GaussianEMFixedPoint.jl) ----------
# Particle velocities are a (3, P) array of rows v[:,1], v[:,2], v[:,3].

def em_irn(c, N):
    return lib().orc_em_irn(float(c), int(N))


def em_stencil(N, c):
    cells, ff = np.zeros(13, dtype=np.int64), np.zeros(13)
    lib().orc_em_stencil(int(N), float(c), cells, ff)
    return cells, ff


def em_frac_bits(P, w, M):
    return lib().orc_em_frac_bits(int(P), float(w), float(M))


def em_deposit(N, y, w, u=None):
    y = _f64(y)
    rho = np.zeros(N)
    uu = None if u is None else _f64(u)
    lib().orc_em_deposit(int(N), y.size, y, _ptr(uu), float(w), rho)
    return rho


def em_fields(eps, dt, rho, Jy, Jz, Fy, Fz, By, Bz, Cy, Cz):
    """One sweep's field updates from the iterate (By, Bz); returns
    (Ex, Ey, Ez, By, Bz, ey, ez, by, bz)."""
    rho, Jy, Jz, Fy, Fz, Cy, Cz = (_f64(a) for a in (rho, Jy, Jz, Fy, Fz, Cy, Cz))
    N = rho.size
    E = np.zeros(3 * N)
    B = np.concatenate([_f64(By), _f64(Bz)])
    e3, b2 = np.zeros(3 * N), np.zeros(2 * N)
    lib().orc_em_fields(N, float(eps), float(dt), rho, Jy, Jz, Fy, Fz, Cy, Cz, E, B, e3, b2)
    return E[:N], E[N:2 * N], E[2 * N:], B[:N], B[N:], e3[N:2 * N], e3[2 * N:], b2[:N], b2[N:]


def em_push(c, ex, ey, ez, by, bz, dt, u):
    ex, ey, ez, by, bz = (_f64(a) for a in (ex, ey, ez, by, bz))
    v = np.zeros(3)
    lib().orc_em_push(ex.size, float(c), ex, ey, ez, by, bz, float(dt), _f64(u), v)
    return v


def em_sort(x, v):
    x = _f64(x).copy()
    v = [_f64(v[k]).copy() for k in range(3)]
    perm = np.zeros(x.size, dtype=np.int64)
    lib().orc_em_sort(x.size, x, v[0], v[1], v[2], perm)
    return x, np.vstack(v), perm


def em_run(N, eps, dt, w, x, v, nsteps, E=None, B=None, Fz=None, rtol=1e-8, max_sweeps=3, sort=1,
           want_fields=True):
    """nsteps of the script's loop.  E: (3, N) Ex, Ey, Ez (zeros); B: (2, N)
    By, Bz (zeros and ones); Fz (zeros).  Returns (x, v, E, B, Fz, fields,
    sweeps), fields of shape (12, nsteps, N) (fields[k, t] = the script's
    fields[:, t, k+1])."""
    x = _f64(x).copy()
    v = [_f64(v[k]).copy() for k in range(3)]
    E = np.zeros(3 * N) if E is None else _f64(E).ravel().copy()
    B = np.concatenate([np.zeros(N), np.ones(N)]) if B is None else _f64(B).ravel().copy()
    Fz = np.zeros(N) if Fz is None else _f64(Fz).copy()
    fields = np.zeros(12 * nsteps * N) if want_fields else None
    sweeps = np.zeros(nsteps, dtype=np.int32)
    rc = lib().orc_em_run(int(N), x.size, float(eps), float(dt), float(w), float(rtol), int(max_sweeps), int(sort),
                          x, v[0], v[1], v[2], E, B, Fz, int(nsteps), _ptr(fields),
                          sweeps.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError(f"orc_em_run rejected its arguments (rc={rc})")
    f = None if fields is None else fields.reshape(12, nsteps, N)
    return x, np.vstack(v), E.reshape(3, N), B.reshape(2, N), Fz, f, sweeps


def emx_gauss_faraday(dt, rho, E, B):
    """src/GaussianEMExplicit.jl:13, :15-16: Ex from rho, (By, Bz) from the
    old (Ey, Ez).  E (3, N), B (2, N).  Returns (E, B, b2) with b2 = (by, bz)."""
    rho = _f64(rho)
    N = rho.size
    E = _f64(E).ravel().copy()
    B = _f64(B).ravel().copy()
    b2 = np.zeros(2 * N)
    lib().orc_emx_gauss_faraday(N, float(dt), rho, E, B, b2)
    return E.reshape(3, N), B.reshape(2, N), b2.reshape(2, N)


def emx_ampere(eps, dt, Jy, Jz, E, B):
    """src/GaussianEMExplicit.jl:23-24: (Ey, Ez) from (By, Bz) and the currents."""
    Jy, Jz = _f64(Jy), _f64(Jz)
    N = Jy.size
    E = _f64(E).ravel().copy()
    lib().orc_emx_ampere(N, float(eps), float(dt), Jy, Jz, E, _f64(B).ravel().copy())
    return E.reshape(3, N)


def emx_push(x, E, b2, dt, u):
    """src/GaussianEMExplicit.jl:17-22 for one particle at x."""
    E, b2 = _f64(E).ravel(), _f64(b2).ravel()
    v = _f64(u).copy()
    lib().orc_emx_push(E.size // 3, float(x), E, b2, float(dt), v)
    return v


def emx_run(N, eps, dt, w, x, v, nsteps, E=None, B=None, sort=1, want_fields=True):
    """nsteps of src/GaussianEMExplicit.jl's loop.  E: (3, N) (zeros); B:
    (2, N) By, Bz (zeros and ones).  Returns (x, v, E, B, fields), fields of
    shape (12, nsteps, N)."""
    x = _f64(x).copy()
    v = [_f64(v[k]).copy() for k in range(3)]
    E = np.zeros(3 * N) if E is None else _f64(E).ravel().copy()
    B = np.concatenate([np.zeros(N), np.ones(N)]) if B is None else _f64(B).ravel().copy()
    fields = np.zeros(12 * nsteps * N) if want_fields else None
    rc = lib().orc_emx_run(int(N), x.size, float(eps), float(dt), float(w), int(sort), x, v[0], v[1], v[2], E, B,
                           int(nsteps), _ptr(fields))
    if rc != 0:
        raise ValueError(f"orc_emx_run rejected its arguments (rc={rc})")
    f = None if fields is None else fields.reshape(12, nsteps, N)
    return x, np.vstack(v), E.reshape(3, N), B.reshape(2, N), f
