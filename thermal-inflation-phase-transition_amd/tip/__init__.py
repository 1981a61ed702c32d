"""tip -- thermal-inflation phase transition lattice engine for MI355X (gfx950).

Host package over libtip.so (HIP kernels behind a C-ABI, include/tip.h):
  tip.kernels          drop-in mirrors of the real-field Numba kernels
  tip.kernels_complex  drop-in mirrors of the complex-field Numba kernels
  tip.potential        V'(phi, T) table producer (host float64)
  tip.device           device arrays over libtip's allocator
"""
from . import _lib  # noqa: F401
from ._lib import TipError, device_count, load  # noqa: F401

__all__ = ["TipError", "device_count", "load"]
__version__ = "0.1.0"
