"""picgolf: MI355X-native electrostatic PIC step of jwscook/ParticleInCellCodeGolf.jl.

The compute runs in libpicgolf.so (HIP kernels for gfx950 behind the C ABI in
include/pic.h, include/pic2d.h, include/pic1d2v.h,
include/pic2d3v.h and include/picem1d.h); this package is the host-side mirror of the scripts'
operators, used from Python (Julia binds the same ABI with ccall).
"""
from ._lib import (LIB_PATH, LibraryMissing, PicError, load, EXPORTS, EXPORTS_2D, EXPORTS_1D2V,  # noqa: F401
                   EXPORTS_2D3V, EXPORTS_EM1D)
from .core import Config, Context  # noqa: F401
from .core2d import Config2D, Context2D  # noqa: F401
from .core1d2v import Config1D2V, Context1D2V  # noqa: F401
from .core2d3v import Config2D3V, Context2D3V, SpeciesConfig  # noqa: F401
from .core_em1d import ConfigEM1D, ContextEM1D  # noqa: F401
from . import configs  # noqa: F401

__version__ = "0.1.0"
