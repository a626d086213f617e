"""ctypes binding of libpicgolf.so (include/pic.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  Loading
fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("PICGOLF_LIB", os.path.join(_PKG, "lib", "libpicgolf.so"))

PIC_OK = 0
PIC_EINVAL = -1
PIC_EHIP = -2
PIC_ENOMEM = -3
PIC_ENONFINITE = -4
PIC_ESTATE = -5

SHAPE_NGP, SHAPE_AREA, SHAPE_GAUSS6, SHAPE_GAUSS7 = 0, 1, 6, 7
RUN_AUTO, RUN_GRAPH, RUN_PERSISTENT = 0, 1, 2
LEAPFROG, MIDPOINT, SIMPSON = 0, 1, 2
FP64, MIXED = 0, 1

# every symbol include/pic.h declares
EXPORTS = [
    "pic_create", "pic_create_batch", "pic_destroy", "pic_last_error", "pic_strerror",
    "pic_version", "pic_upload", "pic_download", "pic_run", "pic_run_profiled",
    "pic_deposit", "pic_fieldsolve", "pic_drift_half", "pic_push_leapfrog",
    "pic_midpoint_position", "pic_push_midpoint", "pic_simpson_kick", "pic_wrap", "pic_isapprox",
    "pic_diagnostics", "pic_set_tiling", "pic_get_tiling", "pic_autotune", "pic_set_run_path",
    "pic_get_run_path", "pic_deposit_noalloc", "pic_noalloc_last_error", "pic_run_async", "pic_sync",
]

# every symbol include/pic2d.h declares
EXPORTS_2D = [
    "pic2d_create", "pic2d_destroy", "pic2d_last_error", "pic2d_upload", "pic2d_download",
    "pic2d_run", "pic2d_deposit", "pic2d_fieldsolve", "pic2d_gather", "pic2d_boris", "pic2d_push",
    "pic2d_diagnostics",
]
PAIR_REFERENCE, PAIR_TEXTBOOK = 0, 1

# every symbol include/pic1d2v.h declares
EXPORTS_1D2V = [
    "pic1d2v_create", "pic1d2v_destroy", "pic1d2v_last_error", "pic1d2v_upload", "pic1d2v_download",
    "pic1d2v_run", "pic1d2v_fields", "pic1d2v_deposit", "pic1d2v_gather", "pic1d2v_boris",
]


# every symbol include/pic2d3v.h declares
EXPORTS_2D3V = [
    "pic2d3v_create", "pic2d3v_destroy", "pic2d3v_last_error", "pic2d3v_upload", "pic2d3v_download",
    "pic2d3v_reset", "pic2d3v_sort", "pic2d3v_run", "pic2d3v_fields", "pic2d3v_field", "pic2d3v_deposit",
    "pic2d3v_gather", "pic2d3v_boris", "pic2d3v_fieldsolve", "pic2d3v_warmup", "pic2d3v_loop", "pic2d3v_lg_get",
    "pic2d3v_lg_set", "pic2d3v_lg_bins", "pic2d3v_lg_grid", "pic2d3v_lg_set_grid", "pic2d3v_lg_iterations",
]
# every symbol include/picem1d.h declares
EXPORTS_EM1D = [
    "picem1d_create", "picem1d_destroy", "picem1d_last_error", "picem1d_upload", "picem1d_download",
    "picem1d_run", "picem1d_deposit", "picem1d_fields", "picem1d_push", "picem1d_explicit_fields",
    "picem1d_explicit_push",
]
MAX_SPECIES_2D3V = 4
NGP_2D3V, AREA_2D3V, BSPLINE0_2D3V = 0, 1, 2
FIELD_ACCUMULATE, FIELD_REPLACE = 0, 1
BORIS_REFERENCE, BORIS_TEXTBOOK = 0, 1


class LibraryMissing(RuntimeError):
    pass


class PicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class pic_config(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32), ("shape", ctypes.c_int32), ("scheme", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("P", ctypes.c_int64), ("W", ctypes.c_double),
        ("w", ctypes.c_double), ("dt", ctypes.c_double), ("rtol", ctypes.c_double),
        ("max_sweeps", ctypes.c_int32), ("device", ctypes.c_int32),
    ]


class pic2d_config(ctypes.Structure):
    _fields_ = [
        ("NX", ctypes.c_int32), ("NY", ctypes.c_int32), ("P", ctypes.c_int64),
        ("precision", ctypes.c_int32), ("pairing", ctypes.c_int32), ("dt", ctypes.c_double),
        ("w", ctypes.c_double), ("B0", ctypes.c_double), ("ns", ctypes.c_int32), ("device", ctypes.c_int32),
        ("sort_every", ctypes.c_int32),
    ]


class pic1d2v_config(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32), ("shape", ctypes.c_int32), ("nspecies", ctypes.c_int32), ("bin", ctypes.c_int32),
        ("P", ctypes.c_int64), ("q", ctypes.c_double * 2), ("q_m", ctypes.c_double * 2),
        ("mass", ctypes.c_double * 2), ("w", ctypes.c_double), ("n0", ctypes.c_double), ("B0", ctypes.c_double),
        ("dt", ctypes.c_double), ("device", ctypes.c_int32), ("precision", ctypes.c_int32),
    ]


class pic2d3v_species(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int64), ("shape", ctypes.c_int32), ("charge", ctypes.c_double), ("mass", ctypes.c_double),
        ("weight", ctypes.c_double),
    ]


class pic2d3v_config(ctypes.Structure):
    _fields_ = [
        ("NX", ctypes.c_int32), ("NY", ctypes.c_int32), ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
        ("dt", ctypes.c_double), ("B0", ctypes.c_double * 3), ("buffer", ctypes.c_int32),
        ("ntskip", ctypes.c_int32), ("ngskip", ctypes.c_int32), ("field_update", ctypes.c_int32),
        ("boris", ctypes.c_int32), ("nspecies", ctypes.c_int32),
        ("species", pic2d3v_species * MAX_SPECIES_2D3V), ("device", ctypes.c_int32), ("field", ctypes.c_int32),
        ("theta", ctypes.c_double), ("theta_source", ctypes.c_double), ("rtol", ctypes.c_double),
        ("maxiters", ctypes.c_int32), ("precision", ctypes.c_int32),
    ]


ELECTROSTATIC_2D3V, LORENZ_STAGGERED_2D3V, LORENZ_2D3V, LORENZ_SEMI_IMPLICIT_2D3V = 0, 1, 2, 3
# Lorenz-gauge state slots and binned fields, in include/pic2d3v.h's order
LG_STATE = ["phi0", "phim", "Ax0", "Ay0", "Az0", "Axm", "Aym", "Azm", "rho0", "Jx0", "Jy0", "Jz0",
            "Ex", "Ey", "Ez", "Bx", "By", "Bz"]
LG_STATE_SI = LG_STATE + ["rhom", "Jxm", "Jym", "Jzm"]   # LorenzGaugeSemiImplicitField
LG_BINS = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Ax", "Ay", "Az", "phi", "rho", "Jx", "Jy", "Jz"]


class picem1d_config(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32), ("max_sweeps", ctypes.c_int32), ("P", ctypes.c_int64), ("eps", ctypes.c_double),
        ("dt", ctypes.c_double), ("w", ctypes.c_double), ("rtol", ctypes.c_double), ("sort", ctypes.c_int32),
        ("device", ctypes.c_int32), ("scheme", ctypes.c_int32),
    ]


PICEM1D_FIXED_POINT = 0
PICEM1D_EXPLICIT = 1


class pic_tuning(ctypes.Structure):
    _fields_ = [
        ("tile_particles", ctypes.c_int64), ("block", ctypes.c_int32), ("n_tiles", ctypes.c_int32),
        ("candidates", ctypes.c_int32), ("ms_per_step", ctypes.c_double), ("entry_ms_per_step", ctypes.c_double),
    ]


class pic_kernel_stat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int64), ("total_ms", ctypes.c_double)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def load():
    """Load libpicgolf.so (raises LibraryMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is missing: build it with "
                             "`python particleincellcodegolf.jl_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    ip = ctypes.POINTER(ctypes.c_int32)
    sig = {
        "pic_create": [ctypes.POINTER(vp), ctypes.POINTER(pic_config)],
        "pic_create_batch": [ctypes.POINTER(vp), ctypes.POINTER(pic_config), i32],
        "pic_destroy": [vp],
        "pic_upload": [vp, i32, vp, vp, vp],
        "pic_download": [vp, i32, vp, vp, vp],
        "pic_run": [vp, i64, vp, vp],
        "pic_run_profiled": [vp, i64, vp, vp, ctypes.POINTER(pic_kernel_stat), ip],
        "pic_deposit": [vp, vp, vp, i64, d, i32, vp],
        "pic_fieldsolve": [vp, vp, vp],
        "pic_drift_half": [vp, vp, vp, i64, d],
        "pic_push_leapfrog": [vp, vp, vp, vp, i64, d, i32],
        "pic_midpoint_position": [vp, vp, vp, vp, vp, i64, d],
        "pic_push_midpoint": [vp, vp, vp, vp, vp, vp, i64, d, i32],
        "pic_simpson_kick": [vp, vp, vp, vp, vp, vp, i64, d, i32],
        "pic_wrap": [vp, vp, i64],
        "pic_isapprox": [vp, vp, vp, i32, d, d, ip],
        "pic_diagnostics": [vp, vp, vp, i64, d, vp],
        "pic_set_tiling": [vp, i64, i32],
        "pic_set_run_path": [vp, i32, i32],
        "pic_run_async": [vp, i64, vp, vp],
        "pic_sync": [vp],
        "pic_deposit_noalloc": [vp, vp, i64, d, i32, i32, vp, vp],
        "pic_get_run_path": [vp, vp, vp],
        "pic_get_tiling": [vp, ctypes.POINTER(i64), ip, ip],
        "pic_autotune": [vp, i32, ctypes.POINTER(pic_tuning)],
    }
    sig.update({
        "pic2d_create": [ctypes.POINTER(vp), ctypes.POINTER(pic2d_config)],
        "pic2d_destroy": [vp],
        "pic2d_upload": [vp, vp, vp, vp, vp, vp, vp, vp],
        "pic2d_download": [vp, vp, vp, vp, vp, vp, vp, vp],
        "pic2d_run": [vp, i64, vp],
        "pic2d_deposit": [vp, vp, vp, i64, d, vp],
        "pic2d_fieldsolve": [vp, vp, vp, vp, vp],
        "pic2d_gather": [vp, vp, vp, vp, vp, i64, vp, vp],
        "pic2d_boris": [vp, vp, vp, vp, vp, vp, i64, d, d],
        "pic2d_push": [vp, vp, vp, vp, vp, vp, i64, vp, vp, d, d],
        "pic2d_diagnostics": [vp, vp, vp, vp, vp, i64, d, vp],
    })
    sig.update({
        "pic1d2v_create": [ctypes.POINTER(vp), ctypes.POINTER(pic1d2v_config)],
        "pic1d2v_destroy": [vp],
        "pic1d2v_upload": [vp, i32, vp, vp, vp],
        "pic1d2v_download": [vp, i32, vp, vp, vp],
        "pic1d2v_run": [vp, i64, vp],
        "pic1d2v_fields": [vp, vp, i64],
        "pic1d2v_deposit": [vp, vp, vp],
        "pic1d2v_gather": [vp, vp, vp, i64, vp],
        "pic1d2v_boris": [vp, vp, vp, vp, i64, d, d, d],
    })
    sig.update({
        "pic2d3v_create": [ctypes.POINTER(vp), ctypes.POINTER(pic2d3v_config)],
        "pic2d3v_destroy": [vp],
        "pic2d3v_upload": [vp, i32, vp],
        "pic2d3v_download": [vp, i32, vp],
        "pic2d3v_reset": [vp],
        "pic2d3v_sort": [vp, i32, d, d, vp],
        "pic2d3v_run": [vp, i64, vp],
        "pic2d3v_fields": [vp, vp, vp, vp, i64],
        "pic2d3v_field": [vp, vp, vp],
        "pic2d3v_deposit": [vp, i32, vp, vp, i64, d, vp],
        "pic2d3v_gather": [vp, i32, vp, vp, vp, vp, i64, vp, vp],
        "pic2d3v_boris": [vp, vp, vp, vp, vp, vp, i64, d],
        "pic2d3v_fieldsolve": [vp, vp, vp, vp, vp],
        "pic2d3v_warmup": [vp],
        "pic2d3v_loop": [vp, i64],
        "pic2d3v_lg_get": [vp, i32, vp],
        "pic2d3v_lg_set": [vp, i32, vp],
        "pic2d3v_lg_bins": [vp, i32, vp, i64],
        "pic2d3v_lg_grid": [vp, vp],
        "pic2d3v_lg_set_grid": [vp, vp],
        "pic2d3v_lg_iterations": [vp, vp, i64],
    })
    sig.update({
        "picem1d_create": [ctypes.POINTER(vp), ctypes.POINTER(picem1d_config)],
        "picem1d_destroy": [vp],
        "picem1d_upload": [vp, vp, vp, vp, vp],
        "picem1d_download": [vp, vp, vp, vp, vp],
        "picem1d_run": [vp, i64, vp, vp],
        "picem1d_deposit": [vp, vp, vp, i64, d, vp],
        "picem1d_fields": [vp] + [vp] * 12,
        "picem1d_push": [vp, vp, vp, vp, vp, i64, vp],
        "picem1d_explicit_fields": [vp] * 7,
        "picem1d_explicit_push": [vp, vp, vp, vp, vp, i64, vp],
    })
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.pic_last_error.argtypes = [vp]
    L.pic_last_error.restype = ctypes.c_char_p
    L.pic_noalloc_last_error.argtypes = []
    L.pic_noalloc_last_error.restype = ctypes.c_char_p
    L.pic2d_last_error.argtypes = [vp]
    L.pic2d_last_error.restype = ctypes.c_char_p
    L.pic1d2v_last_error.argtypes = [vp]
    L.pic1d2v_last_error.restype = ctypes.c_char_p
    L.pic2d3v_last_error.argtypes = [vp]
    L.pic2d3v_last_error.restype = ctypes.c_char_p
    L.picem1d_last_error.argtypes = [vp]
    L.picem1d_last_error.restype = ctypes.c_char_p
    L.pic_strerror.argtypes = [ctypes.c_int]
    L.pic_strerror.restype = ctypes.c_char_p
    L.pic_version.argtypes = []
    L.pic_version.restype = ctypes.c_int
    _lib = L
    return L


def ptr(a):
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
