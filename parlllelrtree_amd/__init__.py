"""MI355X-native (gfx950) R-tree build, spatial join and polygon overlay.

The compute path is libprt.so (HIP kernels behind the C-ABI of
include/prt.h); this package is the host-side mirror of the reference's
operators for Python callers and the tests.
"""
from ._lib import PrtError, lib  # noqa: F401

__version__ = "0.1.0"
