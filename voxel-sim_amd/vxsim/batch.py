"""Batched structured scenes (vx_batch_* of include/vx_solver.h).

A `Batch` is one HBM arena of 512-REV tiles shared by many box scenes of
different sizes. Each `tick()` launches every kernel of the step once for all
running scenes; each scene moves through its own substeps, corotation passes
(src/solver.rs:551-570) and refinement sweeps, with its convergence decided
on the device. Scenes can be admitted while others run and released when
done, which is what the serving loop in `vxsim.serve` builds on.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib
from ._abi import (BatchSceneReport, VX_ERR_TOO_LARGE, VX_SCENE_DONE, VX_SCENE_FAILED, VX_SCENE_FREE,
                   VX_SCENE_RUNNING, c_dp, c_u32p)
from .grid import _report, download_bond_forces, download_bonds, scene_desc, step_params
from .solver import SolverError, check, last_error

FREE, RUNNING, DONE, FAILED = VX_SCENE_FREE, VX_SCENE_RUNNING, VX_SCENE_DONE, VX_SCENE_FAILED
TILE = 512


def bytes_per_row() -> int:
    return int(lib().vx_batch_bytes_per_row())


def rows_for_bytes(budget: int) -> int:
    """Arena rows that fit in `budget` bytes of HBM."""
    return int(budget) // bytes_per_row()


class Batch:
    def __init__(self, capacity_rows: int, max_scenes: int = 256, xcd_chunk_tiles: int = 0,
                 scene_revs_hint: float = None):
        """An arena of `capacity_rows` rows for up to `max_scenes` scenes. With
        `scene_revs_hint` (typical REVs per scene) and no explicit chunk, the
        schedule (XCD interleave, iterations per tick) comes from the schedule
        table (vxsim.tune)."""
        h = ctypes.c_void_p()
        check(lib().vx_batch_create(int(capacity_rows), int(max_scenes), int(xcd_chunk_tiles), ctypes.byref(h)))
        self.handle = h
        self.max_scenes = int(max_scenes)
        self._rows = {}
        if scene_revs_hint is not None and max_scenes > 1:
            from . import tune
            sched = tune.lookup(scene_revs_hint, kind="batch")
            self.set_schedule(0 if xcd_chunk_tiles else sched["chunk"], sched["iters"])

    def set_schedule(self, xcd_chunk_tiles: int = 0, iterations_per_tick: int = 0):
        check(lib().vx_batch_set_schedule(self.handle, int(xcd_chunk_tiles), int(iterations_per_tick)))

    def close(self):
        if getattr(self, "handle", None):
            lib().vx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def free_rows(self):
        """(largest contiguous free range, total free), in rows."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().vx_batch_free_rows(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def admit(self, scene, steps: int, params=None, **kw):
        """Admit `scene` for `steps` fixed steps; None when no contiguous tile
        range (or slot) is free. Other errors raise."""
        p = params if params is not None else step_params(**kw)
        desc, keep = scene_desc(scene)
        slot = ctypes.c_uint32()
        status = lib().vx_batch_admit(self.handle, ctypes.byref(desc), ctypes.byref(p), int(steps), ctypes.byref(slot))
        del keep
        if status == VX_ERR_TOO_LARGE:
            return None
        check(status)
        self._rows[int(slot.value)] = scene.n
        return int(slot.value)

    def tick(self) -> int:
        """One scheduler tick; returns how many admitted scenes still run."""
        running = ctypes.c_uint32()
        check(lib().vx_batch_tick(self.handle, ctypes.byref(running)))
        return int(running.value)

    def run(self, max_ticks: int = 1_000_000) -> int:
        ticks = 0
        while self.tick():
            ticks += 1
            if ticks >= max_ticks:
                raise SolverError(4, "batch still running after %d ticks" % max_ticks)
        return ticks

    def report(self, slot: int) -> dict:
        r = BatchSceneReport()
        check(lib().vx_batch_report(self.handle, int(slot), ctypes.byref(r)))
        return _report(r)

    def states(self) -> np.ndarray:
        """VX_SCENE_* state of every slot after the last tick."""
        out = np.zeros(self.max_scenes, np.uint32)
        check(lib().vx_batch_states(self.handle, out.ctypes.data_as(c_u32p)))
        return out

    def error(self, slot: int):
        """The SolverError of a failed scene, else None."""
        r = self.report(slot)
        if r["state"] != FAILED:
            return None
        return SolverError(r["status"], "scene %d failed with status %d" % (slot, r["status"]))

    def download(self, slot: int) -> dict:
        n = self._rows[int(slot)]
        pos = np.empty((n, 3))
        rot = np.empty((n, 4))
        vel = np.empty((n, 3))
        check(lib().vx_batch_download(self.handle, int(slot), pos.ctypes.data_as(c_dp), rot.ctypes.data_as(c_dp),
                                      vel.ctypes.data_as(c_dp)))
        return dict(positions=pos, rotations=rot, velocities=vel)

    def download_bonds(self, slot: int) -> dict:
        return download_bonds(lib().vx_batch_download_bonds, self.handle, int(slot), n=self._rows[int(slot)])

    def download_bond_forces(self, slot: int) -> dict:
        return download_bond_forces(lib().vx_batch_download_bond_forces, self.handle, int(slot),
                                    n=self._rows[int(slot)])

    def hold(self, slot: int, hold: bool = True):
        """Hold a running scene out of the ticks (or resume it)."""
        check(lib().vx_batch_hold(self.handle, int(slot), 1 if hold else 0))

    def release(self, slot: int):
        check(lib().vx_batch_release(self.handle, int(slot)))
        self._rows.pop(int(slot), None)


__all__ = ["Batch", "bytes_per_row", "rows_for_bytes", "FREE", "RUNNING", "DONE", "FAILED", "TILE", "last_error"]
