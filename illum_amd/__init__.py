"""MI355X-native form-factor / horizon / radiosity / Crank-Nicolson path of illum.

The compute lives in libillum_hip.so (HIP kernels for gfx950 behind the C ABI of
include/illum_hip.h); this package is the host-side binding.
"""
from ._lib import IllumError, load  # noqa: F401
from .api import (  # noqa: F401
    BVH, Context, DeviceArray, IllumContext, Mesh, SparseMatrix, ThermalModel, TimeLoop,
    conductionQ, default_context, fib_spiral, radiosity_solve, synth_ellipsoid)
