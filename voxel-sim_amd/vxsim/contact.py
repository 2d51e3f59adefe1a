"""Voxel contact projection on the device: the host mirror of
`collision::project` (src/collision.rs:33-123) over `vx_contact_project`.

`project(...)` takes the scene fields the reference reads (positions,
previous, rest positions, masses, kinematic flags, per-axis targets and the
bond neighbours) and returns the projected positions and previous positions,
with the contacts sorted by (i, j), as the reference's `Vec<Contact>`.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib
from . import _abi
from .solver import SolverError

EMPTY = 0xFFFFFFFF
_CONTACT_DTYPE = np.dtype([("i", "<u4"), ("j", "<u4"), ("normal", "<f4", (3,)), ("new_impact", "<u4")])


def _f32(a, n, k):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.size != n * k:
        raise ValueError("expected %d values, got %d" % (n * k, a.size))
    return a.reshape(-1).copy()


AUTO = 0     # VX_CONTACT_AUTO: relaxation to each pass's fixed point, windows for a pass beyond its budget
WINDOWS = 1  # VX_CONTACT_WINDOWS: every pass by the window walker


def set_method(method: int, max_rounds: int = 0) -> None:
    """How later calls compute the passes (process-wide; identical results).
    max_rounds: relaxation rounds per call before the passes left go to
    windows (0: the library's default)."""
    L = lib()
    st = L.vx_contact_set_method(int(method), int(max_rounds))
    if st != _abi.VX_OK:
        raise SolverError(st, L.vx_last_error().decode(errors="replace"))


def project(positions, previous, rest_positions, mass, kinematic, bond_neighbor, targets=None, capacity=None):
    """Project voxel pairs apart; returns (positions, previous, contacts, report).

    positions, previous, rest_positions: (n, 3) f32; mass: (n,) f32;
    kinematic: (n,) bool; bond_neighbor: (3, n) u32, EMPTY when absent;
    targets: None or (n, 3) bool (targets[i][a] is Some). contacts is a dict of
    arrays i, j (u32), normal (k, 3) f32, new_impact (bool), sorted by (i, j).
    """
    L = lib()
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    n = pos.shape[0] if pos.ndim == 2 else pos.size // 3
    pos = _f32(pos, n, 3)
    prev = _f32(previous, n, 3)
    rest = _f32(rest_positions, n, 3)
    m = _f32(mass, n, 1)
    kin = np.ascontiguousarray(np.asarray(kinematic).reshape(-1) != 0, dtype=np.uint8)
    bonds = np.ascontiguousarray(bond_neighbor, dtype=np.uint32).reshape(-1)
    if kin.size != n or bonds.size != 3 * n:
        raise ValueError("kinematic needs n entries and bond_neighbor 3*n")
    tgt = None
    if targets is not None:
        t = np.asarray(targets, dtype=bool).reshape(n, 3)
        tgt = np.ascontiguousarray(t[:, 0] * 1 + t[:, 1] * 2 + t[:, 2] * 4, dtype=np.uint8)
    report = _abi.ContactReport()

    def call(cap):
        out = (_abi.Contact * max(cap, 1))()
        p, q = pos.copy(), prev.copy()
        st = L.vx_contact_project(
            n, p.ctypes.data_as(_abi.c_fp), q.ctypes.data_as(_abi.c_fp), rest.ctypes.data_as(_abi.c_fp),
            m.ctypes.data_as(_abi.c_fp), kin.ctypes.data_as(_abi.c_u8p),
            tgt.ctypes.data_as(_abi.c_u8p) if tgt is not None else None, bonds.ctypes.data_as(_abi.c_u32p),
            out, cap, ctypes.byref(report))
        return st, p, q, out

    cap = int(capacity) if capacity is not None else max(16, n)
    st, p, q, out = call(cap)
    if st == _abi.VX_ERR_TOO_LARGE and capacity is None and report.contacts > cap:
        cap = int(report.contacts)
        st, p, q, out = call(cap)
    if st != _abi.VX_OK:
        raise SolverError(st, L.vx_last_error().decode(errors="replace"))
    k = int(report.contacts)
    rec = np.frombuffer(out, dtype=_CONTACT_DTYPE, count=k)
    contacts = {
        "i": rec["i"].copy(),
        "j": rec["j"].copy(),
        "normal": rec["normal"].copy().reshape(k, 3),
        "new_impact": rec["new_impact"] != 0,
    }
    rep = {f: getattr(report, f) for f, _ in _abi.ContactReport._fields_}
    return p.reshape(n, 3), q.reshape(n, 3), contacts, rep
