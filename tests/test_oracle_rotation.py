"""Corotational rotation fit (src/solver.rs:327-417) against the polar
decomposition (scipy.linalg.polar), rigid-motion invariance
(src/solver.rs:1008-1021) and its special cases."""
import math

import numpy as np
import pytest
import scipy.linalg

CELL = 0.01


def quat_to_matrix(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("mode,tol", [(1, 1e-6), (0, 2e-5)])
def test_affine_field_gives_polar_rotation(O, scenes, mode, tol):
    rng = np.random.default_rng(7)
    for trial in range(5):
        s = scenes.uniform_box((3, 3, 3), "iron")
        R = scipy.linalg.expm(np.cross(np.eye(3), rng.normal(scale=0.6, size=3)))
        S = np.eye(3) + 0.15 * (lambda a: (a + a.T) / 2)(rng.normal(size=(3, 3)))
        F = R @ S
        s.positions = (s.positions.astype(np.float64) @ F.T).astype(np.float32)
        os_ = O.OracleScene(s)
        O.fit_rotations(os_, mode)
        Rp, _ = scipy.linalg.polar(F)
        for v in range(s.n):
            np.testing.assert_allclose(quat_to_matrix(os_.rotations[v]), Rp, atol=tol)
        if mode == 1:
            # The centre voxel's own gradient estimate from the stored (f32)
            # positions: column a = (p[+a] - p[-a]) / (2 CELL).
            p = s.positions.astype(np.float64)
            Fc = np.stack([(p[13 + d] - p[13 - d]) / (2 * CELL) for d in (1, 3, 9)], axis=1)
            Rc, _ = scipy.linalg.polar(Fc)
            np.testing.assert_allclose(quat_to_matrix(os_.rotations[13]), Rc, atol=1e-11)


def test_rigid_rotation_has_no_strain(O, scenes):
    # src/solver.rs:1008-1021: rotate a scene by q_z(0.7), fit, and every bond
    # is at its rest length along the bond frame's first axis.
    s = scenes.uniform_box((2, 2, 2), "wood_builtin", grain=2, anisotropy=1.0)
    q = np.array([0, 0, math.sin(0.35), math.cos(0.35)])
    s = scenes.rotate_scene(s, q)
    for mode in (O.F32, O.F64):
        os_ = O.OracleScene(s)
        O.fit_rotations(os_, mode)
        for a in range(3):
            for i in range(s.n):
                j = s.neighbor[a, i]
                if j == 0xFFFFFFFF:
                    continue
                frame = O.bond_frame(os_.rotations[i], os_.rotations[j], a, mode)
                c = frame.T @ (os_.positions[j] - os_.positions[i]) - np.array([CELL, 0, 0])
                assert np.linalg.norm(c) < 1e-6


def test_rest_state_keeps_identity(O, scenes):
    s = scenes.mixed_box((4, 3, 2), seed=5)
    os_ = O.OracleScene(s)
    angle = O.fit_rotations(os_, O.F64)
    assert angle < 1e-12
    np.testing.assert_allclose(os_.rotations, np.tile([0, 0, 0, 1.0], (s.n, 1)), atol=1e-12)


def test_degenerate_fit_is_skipped(O, scenes):
    # det F <= 1e-6: rotations unchanged and no angle reported.
    s = scenes.uniform_box((2, 2, 2), "iron")
    s.positions[:] = s.positions[0]
    q = np.array([0.1, 0.2, 0.3, 0.9], np.float32)
    s.rotations[:] = q / np.linalg.norm(q)
    os_ = O.OracleScene(s)
    before = os_.rotations.copy()
    assert O.fit_rotations(os_, O.F64) == 0.0
    np.testing.assert_array_equal(os_.rotations, before)


def test_kinematic_voxels_are_not_fitted(O, scenes):
    s = scenes.uniform_box((3, 3, 3), "iron")
    s.kinematic[:9] = 1
    R = quat_to_matrix(np.array([0.0, 0.3, 0.0, 0.95]))
    s.positions = (s.positions.astype(np.float64) @ R.T).astype(np.float32)
    os_ = O.OracleScene(s)
    O.fit_rotations(os_, O.F64)
    np.testing.assert_array_equal(os_.rotations[:9], np.tile([0, 0, 0, 1.0], (9, 1)))
    np.testing.assert_allclose(quat_to_matrix(os_.rotations[20]), R, atol=1e-6)


def test_angle_reports_largest_change(O, scenes):
    s = scenes.uniform_box((2, 2, 2), "iron")
    t = 0.2
    s = scenes.rotate_scene(s, np.array([math.sin(t / 2), 0, 0, math.cos(t / 2)]))
    os_ = O.OracleScene(s)
    angle = O.fit_rotations(os_, O.F64)
    assert angle == pytest.approx(t, abs=1e-6)
