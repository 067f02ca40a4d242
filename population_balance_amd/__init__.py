"""MI355X (gfx950) implementation of the lesgo-lineage LES hot path of
adiaiyer/Population_Balance: spectral derivatives, 3/2-dealiased convection,
Lagrangian scale-dependent SGS, batched tridiagonal pressure projection.

The compute path is the HIP library built from csrc/ (liblesgo_hip.so) behind the
C ABI of include/lesgo_amd.h; lesgo.py is the host-side binding.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
