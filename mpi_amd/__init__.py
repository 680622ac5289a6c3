"""MI355X-native DirectPseudoSpectral Helmholtz solver (Fourier x-y, GLL z).

The compute path is libhz.so (hand-written HIP for gfx950 behind the C ABI
of include/hz.h); this package only binds it for Python callers.
"""
from ._lib import HzError, load  # noqa: F401

__all__ = ["HzError", "load"]
