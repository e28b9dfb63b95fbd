"""Host-side mirror of the PUSCH MIMO receive chain over libpuschrx_hip.so.

The compute path is the HIP library; this package only builds configs,
allocates device buffers through torch and enqueues the C ABI entry points.
"""
from . import abi
from .abi import load_library
from .profiles import (FULL_CARRIER_4RX, MID_4RX_RANK2, REFERENCE_FD16X64, SISO_52PRB, Profile)

__all__ = ["abi", "load_library", "Profile", "REFERENCE_FD16X64", "FULL_CARRIER_4RX", "MID_4RX_RANK2",
           "SISO_52PRB"]
