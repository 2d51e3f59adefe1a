"""The rigid tool: the host construction (vxsim.tool.Tool.new, restating
Tool::new) and the oracle's Tool::advance against the reference's tool tests
(src/tool.rs:491-665), in the reference's f32 and in f64:

* the tapered blade's cells are centred on their centroid, keep the blade's
  extent along its body X axis, and the inverse inertia inverts the cells'
  inertia (:495-515);
* a persistent external force accelerates the tool by F/m per unit time and
  stops accelerating once cleared (:517-536);
* an off-centre contact with a free REV conserves linear momentum and angular
  momentum about the original centre of mass, moves the REV and does not add
  kinetic energy (:555-590);
* friction of a spinning tool turns kinetic energy into heat (:592-626);
* a whole step with a one-REV scene leaves no overlap and adds no kinetic
  energy (:628-663)."""
import numpy as np
import pytest

from vxsim import tool as vt

MODES = {"f32": 0, "f64": 1}
WOOD_CHOP_TOOL = (50, 5, 20)  # scenes/wood_chop.ron tool_dimensions


def one_rev(scenes):
    s = scenes.uniform_box((1, 1, 1), "wood_builtin", grain=2, anisotropy=1.0)
    return s


def kinetic(s_vel, s_mass, kin, t):
    wood = sum(0.5 * float(m) * float(v @ v) for v, m, k in zip(s_vel, s_mass, kin) if not k)
    R = _mat(t["rotation"])
    inertia = 1.0 / np.asarray(t["inverse_inertia"])
    w = t["angular_velocity"]
    ang = 0.5 * w @ (R @ (inertia * (R.T @ w)))
    return wood + 0.5 * t["mass"] * float(t["velocity"] @ t["velocity"]) + ang


def _mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _state(ot, tool):
    st = ot.state()
    st["mass"] = tool.mass
    st["inverse_inertia"] = tool.inverse_inertia
    return st


def test_tapered_blade_geometry():
    origin = np.array([0.4, 0.2, 0.3])
    t = vt.Tool.new(vt.BLADE, WOOD_CHOP_TOOL, origin, [0, 1, 0], 0.0)
    assert np.linalg.norm(t.cells.astype(np.float64).mean(axis=0)) < 1e-6
    rel = t.world_cells() - origin
    R = _mat(t.rotation)
    body_x = (rel @ R)[:, 0]
    assert abs(body_x.min() - (-(WOOD_CHOP_TOOL[0] - 1.0) * 0.01 * 0.5)) < 1e-6
    cell_mass = t.mass / len(t.cells)
    c = t.cells.astype(np.float64)
    inertia = ((c * c).sum(axis=1, keepdims=True) + 0.01 * 0.01 / 6.0 - c * c).sum(axis=0) * cell_mass
    assert np.linalg.norm(inertia * t.inverse_inertia - 1.0) < 1e-5


@pytest.mark.parametrize("mode", list(MODES))
def test_external_force_accelerates_and_persists(O, scenes, mode):
    s = one_rev(scenes)
    t = vt.Tool.new(vt.BLADE, (2, 2, 2), [1.0, 1.0, 1.0], [1, 0, 0], 0.0)
    t.external_force = np.array([t.mass * 3.0, 0.0, 0.0])
    ot = O.OracleTool(t)
    vel = s.velocities.astype(np.float64)
    args = (s.positions, vel, s.mass, s.kinematic, None)
    for expect in (0.03, 0.06):
        O.tool_advance(ot, *args, 0.01, [0, 0, 0], MODES[mode])
        assert abs(ot.state()["velocity"][0] - expect) < 1e-6
    for k in range(3):
        ot.struct.external_force[k] = 0.0
    O.tool_advance(ot, *args, 0.01, [0, 0, 0], MODES[mode])
    assert abs(ot.state()["velocity"][0] - 0.06) < 1e-6


def _contact_setup(O, scenes, friction, spin=0.0):
    s = one_rev(scenes)
    t = vt.Tool.new(vt.BLADE, (2, 2, 2), [0.0199, 0.006, 0.005], [-1, 0, 0], 0.1, friction=friction)
    t.angular_velocity = np.array([0.0, 0.0, spin])
    return s, t, O.OracleTool(t)


@pytest.mark.parametrize("mode", list(MODES))
def test_off_center_contact_preserves_momentum(O, scenes, mode):
    s, t, ot = _contact_setup(O, scenes, friction=0.0)
    vel = s.velocities.astype(np.float64)
    m0 = float(s.mass[0])

    def momentum(st):
        p = st["velocity"] * t.mass + vel[0] * m0
        R = _mat(st["rotation"])
        spin = R @ ((1.0 / t.inverse_inertia) * (R.T @ st["angular_velocity"]))
        return p, spin

    before = momentum(_state(ot, t))
    com = ot.state()["position"]
    e0 = kinetic(vel, s.mass, s.kinematic, _state(ot, t))
    O.tool_advance(ot, s.positions, vel, s.mass, s.kinematic, None, 0.0, [0, 0, 0], MODES[mode])
    after = momentum(_state(ot, t))
    orbital = np.cross(s.positions[0].astype(np.float64) - com, vel[0] * m0)
    assert np.linalg.norm(after[0] - before[0]) < 1e-7
    assert np.linalg.norm(after[1] + orbital - before[1]) < 1e-8
    assert np.linalg.norm(vel[0]) > 0.0
    assert kinetic(vel, s.mass, s.kinematic, _state(ot, t)) <= e0 * (1.0 + 1e-5)


@pytest.mark.parametrize("mode", list(MODES))
def test_friction_converts_energy_to_heat(O, scenes, mode):
    s, t, ot = _contact_setup(O, scenes, friction=0.35, spin=10.0)
    vel = s.velocities.astype(np.float64)
    e0 = kinetic(vel, s.mass, s.kinematic, _state(ot, t))
    O.tool_advance(ot, s.positions, vel, s.mass, s.kinematic, None, 0.0, [0, 0, 0], MODES[mode])
    st = _state(ot, t)
    assert st["friction_heat"] > 0.0
    assert kinetic(vel, s.mass, s.kinematic, st) <= e0 * (1.0 + 1e-5)


@pytest.mark.parametrize("mode", list(MODES))
def test_one_cube_step_leaves_no_overlap_and_adds_no_energy(O, scenes, mode):
    s = one_rev(scenes)
    s.config = scenes.StepConfig(fixed_dt=0.001, substeps=1, iterations=40, corotation_iterations=1,
                                 solver_relative_tolerance=1e-5, corotation_tolerance=1e-5, gravity=(0.0, 0.0, 0.0))
    t = vt.Tool.new(vt.BLADE, (2, 2, 2), [0.0199, 0.005, 0.005], [-1, 0, 0], 0.1, friction=0.0)
    ot = O.OracleTool(t)
    os_ = O.OracleScene(s)
    e0 = kinetic(os_.velocities, s.mass, s.kinematic, _state(ot, t))
    O.step_tool(os_, s.config, ot, MODES[mode], O.IC0)
    assert ot.state()["contacts"] > 0
    assert kinetic(os_.velocities, s.mass, s.kinematic, _state(ot, t)) <= e0 * (1.0 + 1e-4)
    # a zero-length advance on the result finds no overlap deeper than 1e-6
    probe = O.OracleTool(t)
    probe.struct = ot.struct
    vel = os_.velocities.copy()
    O.tool_advance(probe, os_.positions, vel, s.mass, s.kinematic, None, 0.0, [0, 0, 0], MODES[mode])
    assert probe.state()["contacts"] == 0


@pytest.mark.parametrize("mode", list(MODES))
def test_fully_fixed_base_stays_put_and_unmount_restores_gravity(O, scenes, mode):
    """tests/supports.rs:81-98: a base mounted with a zero direction holds the
    tool still against a force; unmounted and unloaded, gravity pulls it
    down. (The reference compares positions exactly in f32; the host builds
    the mount's world base in f64, so 1e-7 here.)"""
    s = one_rev(scenes)
    t = vt.Tool.new(vt.BLADE, (50, 5, 20), [1.0, 1.0, 1.0], [1, 0, 0], 3.0)
    t.mount_base([-0.01, 0.0, 0.0], [0.0, 0.0, 0.0])
    t.external_force = np.array([100.0, 100.0, 100.0])
    ot = O.OracleTool(t)
    start = ot.state()["position"].copy()
    vel = s.velocities.astype(np.float64)
    args = (s.positions, vel, s.mass, s.kinematic, None)
    O.tool_advance(ot, *args, 0.05, [0, 0, -9.81], MODES[mode])
    st = ot.state()
    assert np.abs(st["position"] - start).max() < 1e-7
    assert np.abs(st["velocity"]).max() == 0.0
    for k in range(3):
        ot.struct.external_force[k] = 0.0
    ot.struct.has_mount = 0
    O.tool_advance(ot, *args, 0.01, [0, 0, -9.81], MODES[mode])
    st = ot.state()
    assert st["position"][2] < start[2]
    assert st["velocity"][2] < -0.09


@pytest.mark.parametrize("mode", list(MODES))
def test_oblique_guide_preserves_its_line_and_projected_force(O, scenes, mode):
    """tests/supports.rs:113-131: a wedge on a guide along (1, 1, 0)/sqrt(2),
    pushed along X by 4 m for 0.05 s, moves at (0.1, 0.1, 0) and its base
    stays on the guide line."""
    s = one_rev(scenes)
    t = vt.Tool.new(vt.WEDGE, (50, 5, 20), [1.0, 1.0, 1.0], [1, 0, 0], 0.0)
    direction = np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
    base = np.array([-0.01, 0.0, 0.0])
    t.mount_base(base, direction)
    world_base = t.mount[1]
    t.external_force = np.array([t.mass * 4.0, 0.0, 0.0])
    ot = O.OracleTool(t)
    vel = s.velocities.astype(np.float64)
    O.tool_advance(ot, s.positions, vel, s.mass, s.kinematic, None, 0.05, [0, 0, 0], MODES[mode])
    st = ot.state()
    assert np.linalg.norm(st["velocity"] - np.array([0.1, 0.1, 0.0])) < 1e-6
    displacement = st["position"] + vt._qrot64(st["rotation"], base) - world_base
    assert np.linalg.norm(np.cross(displacement, direction)) < 2e-7
    assert np.linalg.norm(displacement) > 1e-3


@pytest.mark.parametrize("shape", [vt.ROD, vt.ROUNDED])
def test_rod_and_rounded_tool_invariants(shape):
    """Tool::new's rod and rounded shapes (src/tool.rs:60-100), which the
    reference's tests do not reach, by invariants that a shifted or transposed
    grid would break: the cells centred on the origin, the mass as the cell
    count x density x CELL^3, the sphere's inertia equal about every axis, and
    the rod's inertia about its axis equal to the cells' distances from it
    (plus each cube's own CELL^2 / 6) with a circular section (equal inertia
    about the two other axes)."""
    length = 12
    t = vt.Tool.new(shape, (length, 5, 20), [0.3, 0.2, 0.1], [1, 0, 0], 0.0)
    c = t.cells.astype(np.float64)
    assert np.abs(c.mean(axis=0)).max() < 1e-6
    density = {vt.ROD: 650.0, vt.ROUNDED: 2700.0}[shape]
    assert t.mass == pytest.approx(len(c) * density * 1e-6, rel=1e-6)
    inertia = 1.0 / t.inverse_inertia
    m = t.mass / len(c)
    if shape == vt.ROUNDED:
        assert len(c) == 1472  # offsets (i + 0.5 - 7) with |p|^2 <= 49
        assert np.ptp(inertia) <= 1e-5 * inertia.max()
    else:
        assert len(c) == 12 * length  # 12 of the 16 section offsets lie within radius 2
        assert np.ptp(c[:, 0]) == pytest.approx((length - 1) * 0.01, rel=1e-6)
        assert inertia[0] == pytest.approx(m * ((c[:, 1] ** 2 + c[:, 2] ** 2).sum() + len(c) * 1e-4 / 6), rel=1e-5)
        assert inertia[1] == pytest.approx(inertia[2], rel=1e-5)
