"""MI355X-native Vertex Block Descent for spg.

Native code: spg_amd/lib/libspg_vbd.so (HIP kernels + C-ABI, include/spg_vbd.h) and
spg_amd/lib/libspg_host.so (C++ mirror of spg's ParticleGroup/energies/BaseSolver, include/spg_host.h).
This package binds them with ctypes; there is no CPU fallback.
"""
from . import native, scenes  # noqa: F401
from .native import ParticleGroup, VBDGpu, VbdContext  # noqa: F401
