"""visondf_amd: MI355X-native (gfx950) radix argsort and sort-based group-by.

The compute path is libvdf_hip.so (HIP kernels behind the C ABI in
include/vdf.h).  `visondf_amd.device.Vdf` drives it over torch device tensors;
include/vison_gpu.hpp is the header-only C++ façade.
"""
from . import _native  # noqa: F401

__all__ = ["_native", "datagen"]
