"""cil_cuda_amd — MI355X (gfx950) overdamped cell-colony MD step behind a C ABI.

The hot path (bin build, pair forces, intra-cell forces, Euler update, division) lives in the
HIP library `libcil_hip.so` built from `csrc/`; this package holds the host-side mirror of
the reference's colony interfaces (`Colony`, `HipMD`) and the workload builders.
"""
from . import abi  # noqa: F401
from .colony import Colony, HipMD  # noqa: F401

__all__ = ["abi", "Colony", "HipMD"]
