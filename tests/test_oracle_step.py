"""End-to-end steps of the oracle's solve driver against the reference's own
step-level tests:

* tests/cpu_preconditioners.rs:8-79  IC(0) and block-Jacobi agree within 2e-5 m
* src/solver.rs:1174-1205             uniform axial traction, max/min < 1.02
* src/solver.rs:1162-1173             equilibrium: no motion at rest
* src/solver.rs:905-932               force and acceleration use the voxel mass
"""
import numpy as np
import pytest

from vxsim.scenes import StepConfig


def permuted(scenes, s, order, O):
    """Reorder voxels into a spawn order; topology from the lattice keys."""
    out = s.copy()
    for name in ("positions", "rotations", "velocities", "mass", "kinematic", "young", "young_perp", "shear",
                 "solid", "fibers", "anisotropy", "delta_temp", "alpha_par", "alpha_perp", "targets", "loads",
                 "keys"):
        setattr(out, name, np.ascontiguousarray(getattr(s, name)[order]))
    out.neighbor, _ = O.build_topology(out.keys, out.kinematic)
    for name in ("damage", "kscale"):
        setattr(out, name, np.ascontiguousarray(getattr(s, name)[:, order]))
    for name in ("formation_offset", "plastic_strain"):
        setattr(out, name, np.ascontiguousarray(getattr(s, name)[:, order]))
    return out


def loaded_lattice(O, scenes):
    # tests/cpu_preconditioners.rs:8-37: 3x2x2 iron, spawned x-outer, one face
    # fixed after construction (bonds kept), voxel 7 fixed along X only, a
    # 3 kN / 2 kN load on voxel 11.
    base = scenes.uniform_box((3, 2, 2), "iron", anisotropy=0.0)
    base.positions = (base.keys.astype(np.float32) * np.float32(0.01) + np.float32(0.005)).astype(np.float32)
    idx = {tuple(k): v for v, k in enumerate(base.keys - base.keys.min(axis=0))}
    order = [idx[(x, y, z)] for x in range(3) for y in range(2) for z in range(2)]
    s = permuted(scenes, base, np.array(order), O)
    s.kinematic[[0, 1, 2, 3]] = 1
    s.targets[7, 0] = s.positions[7, 0]
    s.loads[11] = np.array([0.0, 3000.0, 2000.0], np.float32)
    s.config = StepConfig(fixed_dt=0.001, substeps=1, iterations=256, corotation_iterations=2,
                          solver_relative_tolerance=1e-5, corotation_tolerance=1e-5, gravity=(0.0, 0.0, 0.0))
    return s


def run_lattice(O, s, precond):
    os_ = O.OracleScene(s)
    fixed_x = os_.positions[7, 0]
    start = os_.positions[11].copy()
    for _ in range(8):
        rep = O.step(os_, s.config, O.F32, precond, targets=s.targets, loads=s.loads)
        assert np.all(np.isfinite(os_.positions))
        assert rep["solver_residual"] <= 1e-5
        assert os_.positions[7, 0] == fixed_x
    return os_.positions.copy(), rep["solver_residual"], np.linalg.norm(os_.positions[11] - start)


def test_block_jacobi_and_ic0_agree(O, scenes):
    s = loaded_lattice(O, scenes)
    ic, ic_res, _ = run_lattice(O, s, O.IC0)
    bj, bj_res, moved = run_lattice(O, s, O.BJ)
    assert ic_res <= 1e-5 and bj_res <= 1e-5
    assert moved > 0.0
    assert np.linalg.norm(ic - bj, axis=1).max() < 2e-5


def test_uniform_axial_traction(O, scenes):
    # 10x2x2 wood (grain Z, anisotropy 1), ends targeted, +2e-5 m along X.
    s = scenes.uniform_box((10, 2, 2), "wood_builtin", grain=2, anisotropy=1.0)
    for i in range(s.n):
        if i % 10 == 0:
            s.targets[i] = s.positions[i]
        if i % 10 == 9:
            s.targets[i] = s.positions[i] + np.array([2e-5, 0, 0], np.float32)
    cfg = StepConfig(fixed_dt=0.001, substeps=1, iterations=100, corotation_iterations=4,
                     solver_relative_tolerance=1e-5, corotation_tolerance=1e-5, gravity=(0.0, 0.0, 0.0))
    os_ = O.OracleScene(s)
    rep = O.step(os_, cfg, O.F32, O.IC0, targets=s.targets, want_forces=True)
    fx = rep["last_force"][0, :, 0][s.neighbor[0] != 0xFFFFFFFF]
    assert fx.min() > 0 and fx.max() / fx.min() < 1.02


def test_equilibrium_does_not_move(O, scenes):
    s = scenes.uniform_box((2, 2, 2), "wood_builtin", grain=2, anisotropy=1.0)
    cfg = StepConfig(fixed_dt=0.016666668, substeps=4, iterations=40, corotation_iterations=4,
                     solver_relative_tolerance=1e-5, corotation_tolerance=1e-5, gravity=(0.0, 0.0, 0.0))
    os_ = O.OracleScene(s)
    start = os_.positions.copy()
    O.step(os_, cfg, O.F32, O.IC0)
    assert np.linalg.norm(os_.positions - start, axis=1).max() < 1e-6


def test_force_and_acceleration_use_mass(O, scenes):
    s = scenes.uniform_box((1, 1, 1), "wood_builtin")
    m = float(s.mass[0])
    loads = np.array([[3.0 * m, 2.0 * m, 0.0]], np.float32)
    cfg = StepConfig(fixed_dt=0.01, substeps=1, iterations=1, corotation_iterations=1,
                     solver_relative_tolerance=1e-5, gravity=(0.0, 0.0, 0.0))
    os_ = O.OracleScene(s)
    for k in (1, 2):
        O.step(os_, cfg, O.F32, O.IC0, loads=loads)
        assert abs(os_.velocities[0, 0] - 0.03 * k) < 1e-5
        assert abs(os_.velocities[0, 1] - 0.02 * k) < 1e-5


@pytest.mark.parametrize("mode", [0, 1])
def test_tension_step_reaches_tolerance(O, scenes, mode):
    # The 10^3 tension scene: one 1 ms step, up to four passes, FP64 BJ-PCG.
    s = scenes.tension_10cube()
    os_ = O.OracleScene(s)
    rep = O.step(os_, s.config, mode, O.BJ, exact=(mode == 1), exact_tolerance=1e-10, targets=s.targets)
    assert rep["passes"] >= 1
    top = s.keys[:, 2] == s.keys[:, 2].max()
    assert np.allclose(os_.positions[top, 2], s.targets[top, 2])
    mid = s.keys[:, 2] == 5
    dz = os_.positions[mid, 2] - s.positions[mid, 2]
    assert np.all(dz > 0) and np.all(dz < 1e-4)
