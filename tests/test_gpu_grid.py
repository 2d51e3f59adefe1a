"""GPU parity of the device-resident structured path (rotation fit, bond
assembly, FP32-inner / FP64-refined PCG, step) against the FP64 oracle,
through the C ABI."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EMPTY = 0xFFFFFFFF


def oracle_rows(O, s, h, gravity, mode=1):
    os_ = O.OracleScene(s)
    t = s.targets.astype(np.float64)
    m = ~np.isnan(t)
    kin = s.kinematic != 0
    # predict (src/solver.rs:500-523): v += g h, targeted axes at rest, p += v h
    v = os_.velocities + np.array(gravity) * h
    v[m] = 0.0
    os_.positions[~kin] += v[~kin] * h
    os_.positions[m] = t[m]
    angle = O.fit_rotations(os_, mode)
    sy = O.assemble(os_, mode, h, os_.positions.copy(), s.free_mask(), s.loads)
    return os_, sy, sy.rows(), angle


def device_assembled(s, h, gravity):
    from vxsim import grid
    gs = grid.GridScene(s)
    gs.predict(h, gravity)
    angle = gs.fit_rotations()
    gs.assemble(h)
    return gs, angle


SCENES = {
    "tension10": lambda sc: sc.tension_10cube(),
    "mixed20": lambda sc: sc.large_64cube(dims=(20, 18, 16)),
    "batchbox": lambda sc: sc.batch_boxes(count=3, seed=9)[2],
}


@pytest.mark.parametrize("name", list(SCENES))
def test_assembly_matches_oracle(gpu, O, scenes, name):
    s = SCENES[name](scenes)
    h, g = s.config.fixed_dt, s.config.gravity
    gs, angle = device_assembled(s, h, g)
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, angle_o = oracle_rows(O, s, h, g)
    assert angle == pytest.approx(angle_o, abs=1e-12)
    assert np.array_equal(rows_g["active"], rows_o["active"])
    assert np.array_equal(rows_g["neighbors"], rows_o["neighbors"])  # block indexing bit-exact
    for key in ("blocks", "diagonal", "rhs"):
        scale = np.abs(rows_o[key]).max()
        assert np.abs(rows_g[key] - rows_o[key]).max() <= 1e-12 * scale, key


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("mode", ["mixed", "f64"])
def test_solve_matches_fp64_oracle(gpu, O, scenes, name, mode):
    from vxsim import grid
    s = SCENES[name](scenes)
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    rep = gs.solve(tolerance=1e-10, inner_tolerance=1e-3, mode=grid.MIXED if mode == "mixed" else grid.F64)
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    x_o, _, rel_o = O.solve_exact(rows_o, 1e-13)
    xa = x[rows_g["active"]]
    assert rep["relative_residual"] <= 1e-10
    # <= 1e-6 relative L2 against the FP64 oracle
    assert np.linalg.norm(xa - x_o) / np.linalg.norm(x_o) <= 1e-6
    inactive = np.setdiff1d(np.arange(s.n), rows_g["active"])
    assert np.all(x[inactive] == 0)
    # The device's true residual agrees with the oracle's matvec on the
    # exported system.
    r = rows_o["rhs"] - O.matvec(rows_o, xa)
    assert np.linalg.norm(r) / np.linalg.norm(rows_o["rhs"]) <= 2e-10


def test_rotation_fit_of_rotated_scene(gpu, O, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(6, 5, 4))
    t = 0.4
    s = scenes.rotate_scene(s, np.array([0.0, math.sin(t / 2), 0.0, math.cos(t / 2)]))
    s.kinematic[:] = 0
    gs = grid.GridScene(s)
    a = gs.fit_rotations()
    rot = gs.download()["rotations"]
    gs.close()
    os_ = O.OracleScene(s)
    ao = O.fit_rotations(os_, O.F64)
    assert a == pytest.approx(ao, abs=1e-12) and a == pytest.approx(t, abs=1e-5)
    sign = np.sign(np.sum(rot * os_.rotations, axis=1))[:, None]
    np.testing.assert_allclose(rot * sign, os_.rotations, atol=1e-12)


def test_device_step_matches_oracle_step(gpu, O, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(12, 12, 12))
    cfg = s.config
    gs = grid.GridScene(s)
    sp = grid.solve_params(tolerance=1e-11, inner_tolerance=1e-3)
    for _ in range(3):
        gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=1, corotation_iterations=2,
                corotation_tolerance=cfg.corotation_tolerance, solve=sp)
    out = gs.download()
    gs.close()
    os_ = O.OracleScene(s)
    start = os_.positions.copy()
    for _ in range(3):
        O.step(os_, cfg, O.F64, O.BJ, exact=True, exact_tolerance=1e-12, loads=s.loads)
    disp_o = os_.positions - start
    disp = out["positions"] - start
    assert np.linalg.norm(disp - disp_o) / np.linalg.norm(disp_o) <= 1e-6
    np.testing.assert_allclose(out["velocities"], os_.velocities, atol=1e-6 * np.abs(os_.velocities).max())


def test_solves_are_deterministic_and_warm_start(gpu, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(16, 16, 16))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    r1 = gs.solve(tolerance=1e-9)
    x1 = gs.download()["correction"]
    r2 = gs.solve(tolerance=1e-9)
    x2 = gs.download()["correction"]
    assert r1["inner_iterations"] == r2["inner_iterations"] and np.array_equal(x1, x2)
    r3 = gs.solve(tolerance=1e-9, warm_start=True)
    assert r3["inner_iterations"] == 0 and r3["outer_sweeps"] == 1
    gs.close()


def test_large_scene_residual_property(gpu, O, scenes):
    """Full 64^3 size: the returned correction satisfies the FP64 system the
    oracle assembles, at the requested relative residual."""
    s = scenes.large_64cube()
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    rep = gs.solve(tolerance=1e-9)
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    assert np.array_equal(rows_g["neighbors"], rows_o["neighbors"])
    r = rows_o["rhs"] - O.matvec(rows_o, x[rows_g["active"]])
    assert np.linalg.norm(r) / np.linalg.norm(rows_o["rhs"]) <= 1e-9 * 1.01
    assert rep["relative_residual"] <= 1e-9


def test_large_scene_pre_summed_reductions(gpu, O, scenes):
    """A scene of more than 1024 tiles (96 x 96 x 64 = 1152 tiles) sums its
    inner partials once per kernel instead of in every workgroup; its solve
    still satisfies the oracle's FP64 system, and a warm start from its own
    solution needs no inner iteration."""
    s = scenes.large_64cube(dims=(96, 96, 64))
    assert -(-s.n // 512) > 1024
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    rep = gs.solve(tolerance=1e-9)
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    again = gs.solve(tolerance=1e-9, warm_start=True)
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    r = rows_o["rhs"] - O.matvec(rows_o, x[rows_g["active"]])
    assert np.linalg.norm(r) / np.linalg.norm(rows_o["rhs"]) <= 1e-9 * 1.01
    assert rep["relative_residual"] <= 1e-9 and rep["inner_iterations"] > 0
    assert again["inner_iterations"] == 0 and again["outer_sweeps"] == 1


def test_schedule_validation(gpu, scenes):
    """Iterations per tick must be even (each chunk graph starts on p buffer 0)
    and at most 256; every accepted schedule converges to the tolerance."""
    from vxsim import grid
    from vxsim.solver import SolverError
    s = scenes.large_64cube(dims=(12, 10, 9))
    cfg = s.config
    with grid.GridScene(s, tuned=False) as gs:
        for bad in (7, 258):
            with pytest.raises(SolverError):
                gs.set_schedule(bad)
        for iters in (2, 8, 32):
            gs.set_schedule(iters)
            h = cfg.fixed_dt / cfg.substeps
            gs.predict(h, cfg.gravity)
            gs.fit_rotations()
            gs.assemble(h)
            r = gs.solve(tolerance=1e-9)
            assert r["relative_residual"] <= 1e-9, iters


@pytest.mark.parametrize("dims", [(8, 8, 16400), (7, 9, 16700)])
@pytest.mark.parametrize("precond", ["block_jacobi", "red_black"])
def test_thin_tall_large_scene(gpu, O, scenes, dims, precond):
    """Scenes of more than 1024 tiles take the large-arena kernels (done flag
    tested before the loads, in-tile neighbours staged in LDS); with an x-y
    plane smaller than a tile, z neighbours are inside the tile too. The
    correction satisfies the oracle's FP64 system with either preconditioner."""
    from vxsim import grid
    s = scenes.large_64cube(dims=dims)
    assert -(-s.n // 512) > 1024
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    pc = grid.RED_BLACK if precond == "red_black" else grid.BLOCK_JACOBI
    rep = gs.solve(grid.solve_params(tolerance=1e-9, max_inner=40000, preconditioner=pc))
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    r = rows_o["rhs"] - O.matvec(rows_o, x[rows_g["active"]])
    assert np.linalg.norm(r) / np.linalg.norm(rows_o["rhs"]) <= 1e-9 * 1.01
    assert rep["relative_residual"] <= 1e-9 and rep["inner_iterations"] > 0
