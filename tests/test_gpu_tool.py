"""GPU parity of the rigid tool (Tool::advance, src/tool.rs:192-431, inside the
step as src/solver.rs:530-543 and 668-683 place it, with the zero-time pass
of src/solver.rs:701-706 after the velocities) against the FP64 oracle,
through the C ABI: the tool's pose and velocities and the scene's positions
and velocities after each step, for the reference's one-REV contact cases
(off-centre hit, spinning tool with friction) and for the cutting scene's
blade driven into a wood block, without and with the fracture pass (the
broken-bond set identical up to the 1e-5 exemption)."""
import numpy as np
import pytest

from test_gpu_fracture import EMPTY, exempt
from vxsim import tool as vt
from vxsim.scenes import FractureConfig, HISTORY_FIELDS

pytestmark = pytest.mark.gpu


def _device(s, t, steps, fracture=None, tol=1e-11):
    from vxsim import grid
    cfg = s.config
    sp = grid.solve_params(tolerance=tol, inner_tolerance=1e-3, max_inner=40000)
    states, reps = [], []
    with grid.GridScene(s) as gs:
        gs.set_tool(t)
        for _ in range(steps):
            reps.append(gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                                corotation_iterations=cfg.corotation_iterations,
                                corotation_tolerance=cfg.corotation_tolerance, solve=sp, fracture=fracture))
            states.append(gs.tool_state())
        return states, reps, gs.download(), gs.download_bonds()


def _oracle(O, s, t, steps, fracture=False):
    os_ = O.OracleScene(s)
    ot = O.OracleTool(t)
    hist = O.FractureHistory(s.n)
    states, reps = [], []
    for _ in range(steps):
        reps.append(O.step_tool(os_, s.config, ot, O.F64, O.BJ, exact=True, exact_tolerance=1e-12, loads=s.loads,
                                fracture=(FractureConfig(), hist, s.strength) if fracture else None))
        states.append(ot.state())
        if fracture:
            for a in [os_.damage, os_.plastic_strain] + [getattr(hist, f) for f in HISTORY_FIELDS]:
                a[...] = a.astype(np.float32)
    return states, reps, os_, hist


def _same_tool(sd, so, scale_v, what):
    assert sd["microsteps"] == so["microsteps"], what
    assert sd["contacts"] == so["contacts"], what
    assert np.abs(sd["position"] - so["position"]).max() <= 1e-12 + 1e-9 * scale_v * 1e-3, what
    assert np.abs(sd["velocity"] - so["velocity"]).max() <= 1e-9 * scale_v, what
    assert np.abs(sd["angular_velocity"] - so["angular_velocity"]).max() <= 1e-9 * max(
        np.abs(so["angular_velocity"]).max(), 1.0), what
    q = sd["rotation"] * np.sign(sd["rotation"] @ so["rotation"])
    assert np.abs(q - so["rotation"]).max() <= 1e-12, what
    # the heat sums the tangential work over contacts whose REV velocities
    # come from the step's solve (relative residual 1e-11 on the device)
    assert sd["friction_heat"] == pytest.approx(so["friction_heat"], rel=1e-7, abs=1e-18), what


@pytest.mark.parametrize("case", ["off_centre", "spinning_friction"])
def test_one_rev_contact_matches_oracle(gpu, O, scenes, case):
    s = scenes.uniform_box((1, 1, 1), "wood_builtin", grain=2, anisotropy=1.0)
    s.config = scenes.StepConfig(fixed_dt=0.001, substeps=1, iterations=40, corotation_iterations=1,
                                 solver_relative_tolerance=1e-5, corotation_tolerance=1e-5, gravity=(0.0, 0.0, 0.0))
    friction = 0.0 if case == "off_centre" else 0.35
    t = vt.Tool.new(vt.BLADE, (2, 2, 2), [0.0199, 0.006, 0.005], [-1, 0, 0], 0.1, friction=friction)
    if case == "spinning_friction":
        t.angular_velocity = np.array([0.0, 0.0, 10.0])
    sd, _, out, _ = _device(s, t, 2)
    so, _, os_, _ = _oracle(O, s, t, 2)
    for k in range(2):
        _same_tool(sd[k], so[k], 0.1, (case, k))
    assert sd[0]["contacts"] > 0
    if case == "spinning_friction":
        assert sd[0]["friction_heat"] > 0.0
    np.testing.assert_allclose(out["velocities"], os_.velocities, rtol=0, atol=1e-12)
    np.testing.assert_allclose(out["positions"], os_.positions, rtol=0, atol=1e-14)


def test_blade_strike_matches_oracle(gpu, O, scenes):
    """The cutting scene's blade (tapered, iron) driven at 5 m/s into a
    24x24x16 wood block, four 1 ms steps with six corotation passes, no
    fracture."""
    s = scenes.wood_block(dims=(24, 24, 16))
    t = scenes.blade_tool(s, (10, 2, 8))
    sd, _, out, _ = _device(s, t, 4)
    so, _, os_, _ = _oracle(O, s, t, 4)
    for k in range(4):
        _same_tool(sd[k], so[k], 5.0, ("step", k))
    assert sum(x["contacts"] for x in sd) > 100
    start = s.positions.astype(np.float64)
    disp_o = os_.positions - start
    assert np.linalg.norm(out["positions"] - start - disp_o) <= 1e-6 * np.linalg.norm(disp_o)
    assert np.abs(out["velocities"] - os_.velocities).max() <= 1e-6 * np.abs(os_.velocities).max()


def test_blade_cut_with_tool_matches_oracle(gpu, O, scenes):
    """The blade mounted on its guide and driven at 20 m/s into the block with
    the fracture pass every step (bonds break from the fourth step): the
    tool's state, the broken bonds (up to the 1e-5 exemption) and the
    positions."""
    s = scenes.wood_block(dims=(24, 24, 16))
    t = scenes.blade_tool(s, (10, 2, 8), speed=20.0, mounted=True)
    steps = 5
    sd, reps_d, out, bonds = _device(s, t, steps, fracture=FractureConfig())
    so, reps_o, os_, hist = _oracle(O, s, t, steps, fracture=True)
    differing = exempt(os_, hist, bonds)
    assert sum(r["broken"] for r in reps_o) > 0
    for k in range(steps):
        assert abs(int(reps_d[k]["fracture"]["bonds_broken"]) - reps_o[k]["broken"]) <= differing
        if differing == 0:
            _same_tool(sd[k], so[k], 20.0, ("step", k))
    start = s.positions.astype(np.float64)
    disp_o = os_.positions - start
    assert np.linalg.norm(out["positions"] - start - disp_o) <= 1e-6 * np.linalg.norm(disp_o)


def test_rounded_tool_two_substeps_matches_oracle(gpu, O, scenes):
    """The rounded tool (a 14-REV stone sphere, src/tool.rs:78-84) dropped at
    3 m/s onto a 20x20x12 wood block, two substeps per step (each with its
    advance and zero-time pass), three steps."""
    import dataclasses
    s = scenes.wood_block(dims=(20, 20, 12))
    s.config = dataclasses.replace(s.config, substeps=2)
    top = float(s.keys[:, 2].max() + 1) * 0.01
    c = s.positions.astype(np.float64).mean(axis=0)
    t = vt.Tool.new(vt.ROUNDED, (14, 14, 14), [0, 0, 0], [0, 0, -1], 3.0, friction=0.35)
    low = (t.world_cells() - t.position)[:, 2].min()
    t.position = np.array([c[0] + 0.0013, c[1] - 0.0021, top + 0.005 + 2e-4 - low])
    sd, _, out, _ = _device(s, t, 3)
    so, _, os_, _ = _oracle(O, s, t, 3)
    for k in range(3):
        _same_tool(sd[k], so[k], 3.0, ("step", k))
    assert min(x["contacts"] for x in sd) > 50
    start = s.positions.astype(np.float64)
    disp_o = os_.positions - start
    assert np.linalg.norm(out["positions"] - start - disp_o) <= 1e-6 * np.linalg.norm(disp_o)
    assert np.abs(out["velocities"] - os_.velocities).max() <= 1e-6 * np.abs(os_.velocities).max()


def _guided_jump(scenes, dims, tilt):
    """A blade on a guide along -Z whose body is tilted by `tilt` rad from
    the guide's normal plane, one REV into the side of a wood block: the
    overlap correction moves a guided tool by pen / (n . d) along its guide
    (src/tool.rs:169-177), here many REVs in one step, beyond the candidate
    box the device gathered for the substep's travel."""
    s = scenes.uniform_box(dims, "wood_builtin", grain=2, anisotropy=1.0)
    s.config = scenes.StepConfig(fixed_dt=0.001, substeps=1, iterations=40, corotation_iterations=1,
                                 solver_relative_tolerance=1e-5, corotation_tolerance=1e-5, gravity=(0.0, 0.0, -9.81))
    lo = s.positions.astype(np.float64).min(axis=0)
    hi = s.positions.astype(np.float64).max(axis=0)
    c = 0.5 * (lo + hi)
    d = np.array([-1.0, 0.0, -tilt])
    t = vt.Tool.new(vt.BLADE, (4, 2, 4), [0, 0, 0], d / np.linalg.norm(d), 0.0)
    xmax = (t.world_cells() - t.position)[:, 0].max()
    t.position = np.array([lo[0] - xmax, c[1], c[2]])
    t.mount_base([float(t.cells[:, 0].min()), 0.0, 0.0], [0.0, 0.0, -1.0])
    return s, t


@pytest.mark.parametrize("dims,tilt,cells", [((3, 3, 3), 0.1, 10), ((3, 3, 3), 0.02, 50), ((3, 3, 24), 0.02, 8)])
def test_guided_jump_beyond_the_candidate_box_matches_oracle(gpu, O, scenes, dims, tilt, cells):
    """The device advance notices the tool leaving its candidate box, undoes
    the attempt, gathers the candidates of a box twice as large and runs it
    again (up to 256 REVs of reach): the step equals the oracle's, which bins
    every REV at every microstep."""
    s, t = _guided_jump(scenes, dims, tilt)
    sd, _, out, _ = _device(s, t, 1)
    so, _, os_, _ = _oracle(O, s, t, 1)
    moved = np.linalg.norm(so[0]["position"] - t.position) / 0.01
    assert moved > 0.8 * cells
    _same_tool(sd[0], so[0], max(1.0, float(np.abs(so[0]["velocity"]).max())), "jump")
    start = s.positions.astype(np.float64)
    disp_o = os_.positions - start
    assert np.linalg.norm(out["positions"] - start - disp_o) <= 1e-6 * np.linalg.norm(disp_o)
    assert np.abs(out["velocities"] - os_.velocities).max() <= 1e-6 * np.abs(os_.velocities).max()


def test_tool_arguments_and_removal(gpu, scenes):
    from vxsim import grid
    from vxsim.solver import SolverError
    s = scenes.wood_block(dims=(8, 8, 6))
    t = scenes.blade_tool(s, (4, 2, 4))
    with grid.GridScene(s) as gs:
        with pytest.raises(SolverError):
            gs.tool_state()  # no tool yet
        bad = vt.Tool.new(vt.BLADE, (4, 2, 4), [0, 0, 1], [0, 0, -1], 1.0)
        bad.mass = -1.0
        with pytest.raises(SolverError):
            gs.set_tool(bad)
        bad = scenes.blade_tool(s, (4, 2, 4), mounted=True)
        bad.mount = (bad.mount[0], bad.mount[1], np.array([0.0, 0.0, -2.0]), bad.mount[3])  # not a unit guide
        with pytest.raises(SolverError):
            gs.set_tool(bad)
        gs.set_tool(t)
        p0 = gs.tool_state()["position"]
        assert np.allclose(p0, t.position)
        gs.set_tool(None)
        with pytest.raises(SolverError):
            gs.tool_state()
        # 1e6 m/s needs 250,000 microsteps of 0.4 REV in a 1 ms substep: refused
        # before any microstep runs, and the scene steps again once the tool is sane
        fast = scenes.blade_tool(s, (4, 2, 4), speed=1e6)
        gs.set_tool(fast)
        cfg = s.config
        step = dict(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=1, corotation_iterations=1,
                    corotation_tolerance=cfg.corotation_tolerance)
        with pytest.raises(SolverError) as err:
            gs.step(**step)
        assert err.value.status == 6  # VX_ERR_INVALID_ARGUMENT
        gs.set_tool(t)
        gs.step(**step)
        assert gs.tool_state()["microsteps"] >= 1
