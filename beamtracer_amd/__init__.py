"""beamtracer_amd: beamtracer's particle stage on AMD Instinct MI355X (gfx950).

A ``hip`` kernel backend for the three kernels of the Vlasov iteration's particle stage
(the fused trajectory tracer, the staggered E gather and the space-charge deposit) with the
reference kernels' signatures (:mod:`beamtracer_amd.kernels`), built as the C-ABI library
``libbeamtracer_hip.so`` (``include/beamtracer_hip.h``) and called through ctypes.
"""

from beamtracer_amd._lib import HipLibraryError, build  # noqa: F401
from beamtracer_amd.kernels import KERNEL_NAMES, KERNELS  # noqa: F401

__version__ = "0.1.0"
