"""Bond frame, anisotropic compliance/stiffness and thermal rest length of the
oracle (src/solver.rs:27-95) against closed forms and the reference's own
unit tests (src/solver.rs:1022-1060, 1089-1161)."""
import math

import numpy as np
import pytest

CELL = 0.01
AL = CELL * CELL / CELL  # AREA / CELL


def two_voxels(scenes, material="wood", fiber=(0, 0, 1), anisotropy=1.0, axis=0):
    dims = [1, 1, 1]
    dims[axis] = 2
    s = scenes.uniform_box(tuple(dims), material, grain=2, anisotropy=anisotropy)
    s.fibers[:] = np.asarray(fiber, np.float32)
    return s


def mat32(scenes, name):
    """Material constants as the f32 values the scene stores."""
    m = scenes.MATERIALS[name]
    return type(m)(m.name, *[float(np.float32(getattr(m, f))) for f in
                             ("density", "young", "young_perp", "shear", "alpha_par", "alpha_perp")])


def qz(t):
    return np.array([0, 0, math.sin(t / 2), math.cos(t / 2)])


def rot_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("mode,rtol", [(1, 1e-12), (0, 2e-6)])
def test_isotropic_identical_endpoints(O, scenes, mode, rtol):
    s = two_voxels(scenes, "iron", anisotropy=0.0)
    k = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), mode)
    m = mat32(scenes, "iron")
    np.testing.assert_allclose(k, np.diag([m.young, m.shear, m.shear]) * AL, rtol=rtol, atol=rtol * m.young * AL)


@pytest.mark.parametrize("mode,rtol", [(1, 1e-12), (0, 2e-6)])
def test_fiber_along_bond_adds_shear_coupling(O, scenes, mode, rtol):
    # Sherman-Morrison reduces to 1/(E + 0.2 G a) along the fibre.
    s = two_voxels(scenes, "wood", fiber=(1, 0, 0), anisotropy=1.0)
    m = mat32(scenes, "wood")
    k = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), mode)
    np.testing.assert_allclose(k, np.diag([m.young + 0.2 * m.shear, m.shear, m.shear]) * AL, rtol=rtol,
                               atol=rtol * m.young * AL)


@pytest.mark.parametrize("mode,rtol", [(1, 1e-12), (0, 2e-6)])
def test_fiber_transverse(O, scenes, mode, rtol):
    s = two_voxels(scenes, "wood", fiber=(0, 1, 0), anisotropy=1.0)
    m = mat32(scenes, "wood")
    k = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), mode)
    np.testing.assert_allclose(k, np.diag([m.young_perp, 1.2 * m.shear, m.shear]) * AL, rtol=rtol,
                               atol=rtol * m.young * AL)


def test_general_fiber_matches_stiffness_tensor(O, scenes):
    """Identical endpoints: K = C^-1 * AREA/CELL where C^-1 is the stiffness
    tensor diag(E_n, G, G)*solid + alpha f f^T written out directly (not via
    the Sherman-Morrison compliance)."""
    rng = np.random.default_rng(3)
    m = mat32(scenes, "wood")
    for _ in range(20):
        f = rng.normal(size=3)
        f /= np.linalg.norm(f)
        a = float(rng.uniform(0, 1))
        solid = float(rng.uniform(0.3, 1.0))
        s = two_voxels(scenes, "wood", fiber=f, anisotropy=a)
        s.solid[:] = solid
        f32 = s.fibers[0].astype(np.float64)
        a, solid = float(s.anisotropy[0]), float(s.solid[0])
        tr = m.young + a * (m.young_perp - m.young)
        en = tr + (m.young - tr) * f32[0] ** 2
        stiff = np.diag([en, m.shear, m.shear]) * solid + 0.2 * m.shear * a * np.outer(f32, f32)
        k = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), 1)
        np.testing.assert_allclose(k, stiff * AL, rtol=1e-10, atol=1e-10 * m.young * AL)


def test_series_combination_of_different_endpoints(O, scenes):
    s = two_voxels(scenes, "iron", anisotropy=0.0)
    st = mat32(scenes, "stone")
    s.young[1], s.young_perp[1], s.shear[1] = st.young, st.young_perp, st.shear
    ir = mat32(scenes, "iron")
    k = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), 1)
    # Two half-bonds in series: 2 / (1/E_a + 1/E_b) per diagonal entry.
    expect = np.diag([2 / (1 / ir.young + 1 / st.young), 2 / (1 / ir.shear + 1 / st.shear),
                      2 / (1 / ir.shear + 1 / st.shear)]) * AL
    np.testing.assert_allclose(k, expect, rtol=1e-12, atol=1e-3)


def test_invalid_compliance_gives_zero(O, scenes):
    s = two_voxels(scenes, "iron", anisotropy=0.0)
    s.shear[1] = 0.0
    k = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), 1)
    assert np.all(k == 0.0)


def test_stiffness_symmetric_and_rotation_invariant(O, scenes):
    # src/solver.rs:1022-1031 (symmetry, shear present, det > 0), and
    # rotating the voxels together with the frame leaves K unchanged.
    s = two_voxels(scenes, "wood", fiber=(0.3, 0.5, 0.81), anisotropy=0.7)
    k = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), 0)
    assert k[1, 1] > 0 and k[2, 2] > 0 and np.linalg.det(k) > 0
    assert np.abs(k - k.T).max() < 1e-3
    q = np.array([0.2, -0.4, 0.1, 0.8])
    q /= np.linalg.norm(q)
    s2 = s.copy()
    s2.rotations[:] = q.astype(np.float32)
    frame = rot_matrix(s2.rotations[0].astype(np.float64))
    k2 = O.stiffness_matrix(O.OracleScene(s2), 0, 1, frame, 1)
    k1 = O.stiffness_matrix(O.OracleScene(s), 0, 1, np.eye(3), 1)
    np.testing.assert_allclose(k2, k1, rtol=1e-6, atol=1e-6 * np.abs(k1).max())


def test_vertical_wood_bonds_stiffer(O, scenes):
    # src/solver.rs:1053-1059: grain-Z wood, vertical / horizontal axial > 15.
    s = scenes.uniform_box((2, 2, 2), "wood_builtin", grain=2, anisotropy=1.0)
    os_ = O.OracleScene(s)
    v = O.stiffness_matrix(os_, 0, int(s.neighbor[2, 0]), O.bond_frame([0, 0, 0, 1], [0, 0, 0, 1], 2), 0)[0, 0]
    h = O.stiffness_matrix(os_, 0, int(s.neighbor[0, 0]), O.bond_frame([0, 0, 0, 1], [0, 0, 0, 1], 0), 0)[0, 0]
    assert v > 15 * h


def test_bond_frame_basis_and_midpoint(O):
    ident = [0, 0, 0, 1]
    np.testing.assert_allclose(O.bond_frame(ident, ident, 1), np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]]), atol=0)
    for t0, t1 in [(0.1, 0.5), (-1.0, 2.0), (0.3, 0.3001)]:
        f = O.bond_frame(qz(t0), qz(t1), 0)
        np.testing.assert_allclose(f, rot_matrix(qz((t0 + t1) / 2)), atol=1e-12)
    # Opposite-sign quaternions describe the same rotation: the shortest arc.
    f = O.bond_frame(qz(0.2), -qz(0.6), 0)
    np.testing.assert_allclose(f, rot_matrix(qz(0.4)), atol=1e-12)


def test_thermal_rest_isotropic(O, scenes):
    # src/solver.rs:1105-1114: iron, alpha 1e-5, +100 K -> rest = 1.001 CELL.
    s = two_voxels(scenes, "iron", anisotropy=0.0)
    s.alpha_par[:] = 1e-5
    s.alpha_perp[:] = 1e-5
    s.delta_temp[:] = 100.0
    for mode in (0, 1):
        rest = O.bond_thermal_rest(O.OracleScene(s), 0, 1, np.eye(3), mode)
        assert abs(rest[0] / CELL - 1.001) < 1e-6
        assert abs(rest[1]) < 1e-12 and abs(rest[2]) < 1e-12


def test_thermal_rest_anisotropic(O, scenes):
    # src/solver.rs:1148-1160: wood 3e-6 / 30e-6, +100 K: across the grain
    # 1.003, along it 1.0003.
    s = two_voxels(scenes, "wood", fiber=(0, 0, 1), anisotropy=1.0, axis=0)
    s.delta_temp[:] = 100.0
    t = O.bond_thermal_rest(O.OracleScene(s), 0, 1, O.bond_frame([0, 0, 0, 1], [0, 0, 0, 1], 0), 1)[0]
    s2 = two_voxels(scenes, "wood", fiber=(0, 0, 1), anisotropy=1.0, axis=2)
    s2.delta_temp[:] = 100.0
    p = O.bond_thermal_rest(O.OracleScene(s2), 0, 1, O.bond_frame([0, 0, 0, 1], [0, 0, 0, 1], 2), 1)[0]
    assert abs(t / CELL - 1.003) < 1e-6
    assert abs(p / CELL - 1.0003) < 1e-6
    assert t > p * 1.002
