"""kelvin_amd: MI355X-native finite-temperature CCSD amplitude engine.

User API mirrors the reference (``kelvin.ccsd.ccsd(sys, T=, mu=, ngrid=)``
with ``.run()``); the amplitude iteration runs in libftcc.so (HIP, gfx950).
"""
from ._lib import FTCCError, LIB_PATH  # noqa: F401


def __getattr__(name):
    # lazy imports keep ``import kelvin_amd`` cheap (no torch import)
    if name == "ccsd":
        from .ccsd import ccsd
        return ccsd
    if name in ("UEGSystem", "PUEGSystem", "UEGBasis", "ueg_from_config"):
        from . import ueg_system
        return getattr(ueg_system, name)
    if name in ("DFModelSystem", "ArraySystem"):
        from . import model_systems
        return getattr(model_systems, name)
    if name == "FTEngine":
        from .engine import FTEngine
        return FTEngine
    raise AttributeError(name)
