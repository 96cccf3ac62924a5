"""MI355X-native solve path for the statically condensed HHO face system.

The product is the shared library ``libfhgpu.so`` (HIP kernels for gfx950,
C++ host Solver API, C ABI in include/fhgpu.h); ``fhhos4_amd.fhgpu`` binds it
with ctypes.
"""
from . import fhgpu  # noqa: F401
from .fhgpu import BF16, F32, F64, Context, FhError, Matrix, Solver  # noqa: F401
