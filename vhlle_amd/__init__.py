"""vhlle_amd: MI355X-native implementation of vHLLE's hydro step.

The product is the native library libvhlle_gpu.so (HIP kernels for gfx950,
the C-ABI declared in include/vhlle_gpu.h and the C++ facade with the
reference's Fluid/Hydro/Cell/EoS/TransportCoeff classes).  This Python package
only locates and binds it (vhlle_amd.abi) and holds the host-side setup helpers
used by tests and the benchmark (vhlle_amd.setups).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
