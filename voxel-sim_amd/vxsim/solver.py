"""Python mirror of the reference's device-solve boundary.

`solve_pcg` has the argument meaning, results and error behaviour of
`cuda_solver::solve_pcg` (src/cuda_solver.rs:301-501): it returns
`(correction, iterations, relative_residual)` or raises `SolverError`
carrying the message the library reports, where the reference returns
`Err(String)`. The work runs in libvxsim.so (vx_solve_pcg).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib
from ._abi import c_fp, c_u32p

MAX_U32 = 0xFFFFFFFF


class SolverError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def last_error() -> str:
    msg = lib().vx_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status != 0:
        raise SolverError(status, last_error())


def _f32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(shape)


def _call(fn, rhs, neighbors, diagonal, blocks, inverse_diagonal, max_iterations, tolerance):
    rhs = np.asarray(rhs)
    n = int(rhs.shape[0]) if rhs.ndim > 1 else int(rhs.size // 3)
    if n > MAX_U32:
        # src/cuda_solver.rs:314-315
        raise SolverError(2, "HIP PCG supports at most u32::MAX active REV")
    rhs = _f32(rhs, (n, 3))
    nbr = np.ascontiguousarray(neighbors, dtype=np.uint32).reshape(n, 6)
    diag = _f32(diagonal, (n, 9))
    blk = _f32(blocks, (n, 54))
    inv = _f32(inverse_diagonal, (n, 9))
    out = np.zeros((n, 3), np.float32)
    it = ctypes.c_uint32(0)
    rel = ctypes.c_float(0.0)
    status = fn(rhs.ctypes.data_as(c_fp), nbr.ctypes.data_as(c_u32p), diag.ctypes.data_as(c_fp),
                blk.ctypes.data_as(c_fp), inv.ctypes.data_as(c_fp), n, int(min(max_iterations, MAX_U32)),
                float(tolerance), out.ctypes.data_as(c_fp), ctypes.byref(it), ctypes.byref(rel))
    check(status)
    return out, int(it.value), float(rel.value)


def solve_pcg(rhs, neighbors, diagonal, blocks, inverse_diagonal, max_iterations, tolerance):
    """Block-Jacobi PCG of the reference device solve (f32 vectors, f64 dots)."""
    return _call(lib().vx_solve_pcg, rhs, neighbors, diagonal, blocks, inverse_diagonal, max_iterations, tolerance)


def solve_pcg_f64(rhs, neighbors, diagonal, blocks, inverse_diagonal, max_iterations, tolerance):
    """The plain FP64 GPU solve over the same inputs."""
    return _call(lib().vx_solve_pcg_f64, rhs, neighbors, diagonal, blocks, inverse_diagonal, max_iterations,
                 tolerance)


def solve_pcg_noalloc(rhs, neighbors, diagonal, blocks, inverse_diagonal, max_iterations, tolerance):
    """solve_pcg without device allocation (vx_solve_pcg_noalloc): the same
    results, at most about 90,000 rows (the library's static scratch area)."""
    return _call(lib().vx_solve_pcg_noalloc, rhs, neighbors, diagonal, blocks, inverse_diagonal, max_iterations,
                 tolerance)
