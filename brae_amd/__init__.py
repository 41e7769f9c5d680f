"""brae_amd: an MI355X-native finite-volume assembly and pressure-solve engine.

The product is the C-ABI shared library libmfv.so (include/mfv.h) built from csrc/ for gfx950;
brae_amd.mfv is its ctypes binding."""
from . import mfv  # noqa: F401
