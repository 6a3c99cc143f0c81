"""dmd_amd: MI355X-native (gfx950) step engine for DMD.

Host-side mirror of DMD's ForceComponent / ForceEngine / SimulationEngine
interface over the C-ABI of dmd_amd/libdmdx.so (include/dmdx.h).
"""
from . import _abi
from .system import SystemSpec

__all__ = ["SystemSpec", "_abi"]
