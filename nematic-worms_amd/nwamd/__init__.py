"""nwamd -- Python side of the MI355X-native Nematic-Worms time step.

The compute lives in libnwamd.so (HIP kernels for gfx950 behind the C-ABI of
include/nw_abi.h); the host driver that sequences it (ParticleSystem,
simulation/particleSystem.cpp) lives in libnwhost.so.  This package only
binds them.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
