"""MI355X-native FMS (full multiple scattering) stage for refeff.

The compute lives in ``librefeff_fms.so`` (HIP kernels for gfx950 behind the C
ABI in ``include/refeff_fms.h``); this package is the thin host-side mirror
of refeff's FMS interface.
"""
from ._lib import FmsError, FmsLibraryMissing, LIB_PATH, header_functions, lib  # noqa: F401

__version__ = "0.1.0"
