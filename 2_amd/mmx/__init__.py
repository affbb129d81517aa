"""mmx -- Python front end of the gfx950 micromagnetic solver core (libmmx.so).

This package is a thin ctypes layer over the C ABI in include/mmx.h; all
computation happens in the HIP library.  There is no CPU fallback: importing
the package without a built libmmx.so raises, and every call that needs the
GPU fails with the library's own error when no GPU is present.

Two levels, as in the reference:
  * `lib` -- the raw C entry points (drop-in kernels, plans, reductor, sim);
  * `Sim` -- the engine object for one simulation (internal z,y,x order), and
    `mmx.api` -- the user-facing script API of mumax/2 (setgridsize, load,
    setv, setarray, run, getv ... in user x,y,z order).
"""
from ._lib import MMXError, lib, load_library, check, LIB_PATH, EXPORTED_SYMBOLS  # noqa: F401
from ._lib import MMX_KERNEL_ARNE, MMX_KERNEL_NEWELL  # noqa: F401
from .sim import Sim, DemagPlan, Reductor, Batch, Surrogate  # noqa: F401
from . import kernels  # noqa: F401

__all__ = ["MMXError", "lib", "load_library", "check", "Sim", "DemagPlan", "Reductor", "Batch", "Surrogate",
           "kernels", "LIB_PATH", "EXPORTED_SYMBOLS", "MMX_KERNEL_ARNE", "MMX_KERNEL_NEWELL"]
