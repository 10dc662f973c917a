"""MI355X-native pressure-correction time step of CaNS (one GPU).

The compute path is ``libcans_hip.so`` (HIP kernels for gfx950 behind the C ABI
declared in include/cans_hip.h); this package is its Python host.
"""
from ._lib import CansError, lib  # noqa: F401
