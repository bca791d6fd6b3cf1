"""nonenone_amd — MI355X-native per-cell flash and Natural assembly (FP64, HIP).

Host mirror of the reference's EOS-flash and Natural-assembly surfaces over
the C ABI in include/nonenone_amd.h.  All compute runs in HIP kernels from the
in-tree libnonenone_amd.so; importing this package does not load it, the first
Context does (and raises if it is missing).
"""
from . import _abi, fluids  # noqa: F401
from ._abi import FlashOptions, FluidDesc  # noqa: F401

__all__ = ["FlashOptions", "FluidDesc", "fluids", "ThreePhaseFlash", "NaturalAssembly"]


def __getattr__(name):
    if name == "ThreePhaseFlash":
        from .flash import ThreePhaseFlash
        return ThreePhaseFlash
    if name == "NaturalAssembly":
        from .assembly import NaturalAssembly
        return NaturalAssembly
    raise AttributeError(name)
