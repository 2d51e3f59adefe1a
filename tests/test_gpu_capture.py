"""The capture tool on the device: the CSV written from `vx_scene_step` with
the fracture pass holds the same positions (as f32), broken-bond totals and
cluster counts as the same scene stepped directly, and the command line takes
the reference's SceneConfig RON file for its step settings."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_capture_matches_direct_steps(gpu, tmp_path):
    from vxsim import capture, grid, interop
    cfg = capture.default_config()
    scene, _ = capture.make_scene("pulled_mixed")
    path = str(tmp_path / "cap.csv")
    totals = capture.run_capture(scene, path, 3, 2, cfg.step_params())
    cap = interop.read_capture(path)
    assert sorted(cap["frames"]) == [0, 3]
    with grid.GridScene(scene) as gs:
        broken = 0
        for _ in range(3):
            broken += gs.step_with(cfg.step_params())["fracture"]["bonds_broken"]
        out, bonds = gs.download(), gs.download_bonds()
    assert broken == totals["bonds_broken"] > 0
    f3 = cap["frames"][3]
    assert f3.broken == broken and f3.clusters == interop.cluster_count(bonds["neighbor"], scene.kinematic)
    idx = f3.samples[:, 6].astype(np.int64)
    assert np.array_equal(f3.samples[:, 1:4].astype(np.float32), out["positions"][idx].astype(np.float32))
    dmg, lost = interop.surface_state(bonds["neighbor"], bonds["damage"], scene.dims, idx)
    assert np.array_equal(f3.samples[:, 4].astype(np.float32), dmg) and np.array_equal(f3.samples[:, 5], lost)
    f0 = cap["frames"][0]
    assert f0.broken == 0 and np.array_equal(f0.samples[:, 1:4].astype(np.float32), scene.positions[idx])


def test_capture_command_line_with_ron_config(gpu, tmp_path):
    from vxsim import capture, interop
    path = str(tmp_path / "cut.csv")
    assert capture.main([path, "2", "2", "--scene", "cut", "--config", os.path.join(GOLDEN, "mounted_axe.ron")]) == 0
    cap = interop.read_capture(path)
    assert cap["header"]["steps"] == 2 and cap["header"]["dt"] == 0.001 and sorted(cap["frames"]) == [0, 2]
    assert cap["frames"][2].broken > 0
    assert capture.main([path, "2", "4"]) == 2  # saved frames beyond steps + 1
