"""MI355X-native AFAR SSP-RK3 pipeline (2D compressible Euler, unstructured FV).

The compute path is libafar.so (hand-written HIP for gfx950 behind the C-ABI
of include/afar.h); this package is its host-side mirror of the reference
solver's interfaces.  Importing does not load the library; the first call
does, and fails loudly if it has not been built.
"""
from ._lib import AfarError, lib  # noqa: F401
from .solver import BasicCUDASolver, DeviceFields, Mesh, device_info, field_pointer, schemes  # noqa: F401

__all__ = ["AfarError", "BasicCUDASolver", "DeviceFields", "Mesh", "device_info", "field_pointer", "lib",
           "schemes"]
