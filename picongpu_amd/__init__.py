"""MI355X-native PIC-cycle hot path with PIConGPU's functor surface.

The compute path is the HIP library built from picongpu_amd/csrc (gfx950),
reached through the C ABI in include/picongpu_amd.h; `abi.Simulation` is the
Python host binding and `setups` holds the .param-style configurations.
"""
from . import abi, setups  # noqa: F401
from .abi import Simulation  # noqa: F401

__all__ = ["abi", "setups", "Simulation"]
