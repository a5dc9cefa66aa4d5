"""MI355X-native pressure projection for SIMCOFlow's two-phase cut-cell step.

The engine (``libsimcoproj.so``) is a HIP/C++ library with a C ABI
(``include/simco_proj.h``).  This package holds its sources, the ctypes
binding used by Python callers and the torch.autograd adjoint wrapper.
"""

from ._paths import LIB_DIR, PKG_DIR  # noqa: F401

__all__ = ["LIB_DIR", "PKG_DIR"]
