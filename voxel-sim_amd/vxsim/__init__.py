"""Host side of the MI355X corotational elastic PCG solver.

`vxsim.solver` mirrors the reference's device boundary
(`cuda_solver::solve_pcg`, src/cuda_solver.rs:301-309) over the C ABI of
include/vx_solver.h; `vxsim.grid` drives the device-resident structured
scenes; `vxsim.scenes` builds the synthetic workloads.

The compute path is the HIP library libvxsim.so next to this file. There is
no CPU fallback: when the library is missing, `lib()` raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvxsim.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load libvxsim.so (built by `make -C voxel-sim_amd/csrc`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                "libvxsim.so is not built (%s); run __graft_entry__.build() or make -C voxel-sim_amd/csrc" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH)
        from . import _abi
        _abi.declare(_lib)
    return _lib
