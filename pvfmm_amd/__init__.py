"""MI355X-native particle FMM with the interface of dmalhotra/pvfmm's particle path.

The compute runs in libpvfmm.so (hand-written gfx950 HIP kernels behind a C ABI); this package
is the Python binding.  There is no CPU fallback: without the built library every call raises
``pvfmm_amd.ffi.LibraryNotBuilt``.
"""
from .wrapper import (  # noqa: F401
    KERNEL_DIMS,
    SUPPORTED_KERNELS,
    FMMKernel,
    FMMParticleContext,
    PtFMM,
    PtFMMBatch,
    PtFMMTree,
    device_count,
    direct,
    kernel_dims,
    kernel_id,
)

__all__ = [
    "FMMKernel",
    "FMMParticleContext",
    "KERNEL_DIMS",
    "SUPPORTED_KERNELS",
    "PtFMM",
    "PtFMMBatch",
    "PtFMMTree",
    "device_count",
    "direct",
    "kernel_dims",
    "kernel_id",
]

__version__ = "0.1.0"
