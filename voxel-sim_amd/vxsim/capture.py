"""Capture a device-resident scene in the reference's H/F/V CSV format.

    python -m vxsim.capture OUT.csv [STEPS] [SAVED_FRAMES] [--scene NAME] [--config FILE.ron]

The arguments and checks follow src/bin/solver_scale_capture.rs:47-62
(STEPS default 176, SAVED_FRAMES default STEPS + 1 and within 2..=STEPS+1);
the records follow its :113-149. Each step is one `vx_scene_step` with the
fracture pass; the step settings come from FILE.ron (a SceneConfig) or the
reference's SceneConfig defaults with a 1 ms step (:63-66). The reference's
own capture scene, an iron block resting on a spawned-fixed bottom layer,
is held by contact, which this library does not model; NAME picks one of
this package's scenes instead (pulled_mixed: a 24x24x16 mixed box pulled
from its top face; cut: the wood block cut by the blade at 48x48x24).
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import grid, interop, scenes


def default_config() -> interop.SceneConfig:
    """SceneConfig::default() (src/scene.rs:244-266) with a 1 ms step."""
    fr = interop.FractureSettings(35.0, 0.005, 1, 1000.0, 3.0, 0.5, 0.9, 10, 10, 0.35, 0.1)
    return interop.SceneConfig(dimensions=(1, 1, 1), branch_dimensions=(0, 0, 0), ground_dimensions=(1, 1, 1),
                               tool_dimensions=(2, 2, 2), fibers=(0.0, 0.0, 1.0), anisotropy=0.0,
                               gravity=(0.0, 0.0, float(np.float32(-9.81))), fixed_dt=float(np.float32(0.001)),
                               substeps=4, iterations=80, solver_relative_tolerance=float(np.float32(1e-5)),
                               corotation_tolerance=float(np.float32(1e-5)), corotation_iterations=2, fracture=fr,
                               initial_speed=0.0, tool_start=(0.0, 0.0, 0.0))


def make_scene(name: str):
    """(scene, per-step targets function or None)."""
    if name == "pulled_mixed":
        return scenes.pulled_mixed((24, 24, 16), 7, 4e-4), None
    if name == "cut":
        blade = (12, 3, 6)
        s = scenes.wood_block_blade(dims=(48, 48, 24), blade=blade)
        return s, (lambda k: scenes.blade_targets(s, blade, 5.0, s.config.fixed_dt, k))
    raise ValueError("unknown scene %r" % name)


def run_capture(scene, path: str, steps: int, saved_frames: int, params, targets_fn=None) -> dict:
    """Step `scene` `steps` times with `params` and write the capture; returns
    the totals (bonds broken, solves, inner iterations)."""
    samples = interop.surface_samples(scene.dims)
    totals = {"bonds_broken": 0, "solves": 0, "inner_iterations": 0}
    with grid.GridScene(scene) as gs, interop.CaptureWriter(path, params.fixed_dt, steps, saved_frames) as cw:
        for frame in range(steps + 1):
            if frame > 0:
                if targets_fn is not None:
                    gs.set_targets_and_loads(targets_fn(frame - 1), scene.loads)
                r = gs.step_with(params)
                totals["bonds_broken"] += int(r["fracture"]["bonds_broken"])
                totals["solves"] += r["solves"]
                totals["inner_iterations"] += r["inner_iterations"]
            if cw.saves(frame):
                out, bonds = gs.download(), gs.download_bonds()
                dmg, lost = interop.surface_state(bonds["neighbor"], bonds["damage"], scene.dims, samples[:, 1])
                cw.frame(frame, totals["bonds_broken"], interop.cluster_count(bonds["neighbor"], scene.kinematic),
                         samples, out["positions"], dmg, lost)
    return totals


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m vxsim.capture")
    ap.add_argument("path", nargs="?", default="solver-scale.csv")
    ap.add_argument("steps", nargs="?", type=int, default=176)
    ap.add_argument("saved_frames", nargs="?", type=int, default=None)
    ap.add_argument("--scene", default="pulled_mixed", choices=("pulled_mixed", "cut"))
    ap.add_argument("--config", default=None, help="SceneConfig RON file for the step settings")
    a = ap.parse_args(argv)
    saved = a.steps + 1 if a.saved_frames is None else a.saved_frames
    if a.steps == 0 or saved < 2 or saved > a.steps + 1:
        print("steps must be positive and saved_frames must be in 2..=steps+1", file=sys.stderr)
        return 2
    cfg = interop.load_scene_config(a.config) if a.config else default_config()
    scene, targets_fn = make_scene(a.scene)
    totals = run_capture(scene, a.path, a.steps, saved, cfg.step_params(), targets_fn)
    print("capture: %d REV, %d steps, %d states, %d bonds broken, %d solves, %d inner iterations -> %s" %
          (scene.n, a.steps, saved, totals["bonds_broken"], totals["solves"], totals["inner_iterations"], a.path),
          file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
