"""Measurement aids change nothing in the results: the PCG chunks launched
kernel by kernel (vx_set_launch_mode) and the dispatch-timed solve
(vx_scene_solve_timed) give the same bits as the captured-graph replay, for
the structured solve, a whole step and the drop-in row solve."""
import numpy as np
import pytest

from test_gpu_grid import device_assembled

pytestmark = pytest.mark.gpu


def _solve_and_step(scenes, direct):
    from vxsim import grid
    grid.set_launch_mode(direct)
    try:
        s = scenes.large_64cube(dims=(20, 18, 16))
        cfg = s.config
        gs, _ = device_assembled(s, cfg.fixed_dt, cfg.gravity)
        rep = gs.solve(tolerance=1e-10, inner_tolerance=3e-3)
        x = gs.download()["correction"]
        gs.close()
        with grid.GridScene(s) as gs:
            srep = gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, corotation_iterations=2,
                           solve=grid.solve_params(tolerance=1e-9))
            out = gs.download()
    finally:
        grid.set_launch_mode(False)
    return rep, x, srep, out


def test_direct_launches_equal_graph_replay(gpu, scenes):
    rep_g, x_g, srep_g, out_g = _solve_and_step(scenes, False)
    rep_d, x_d, srep_d, out_d = _solve_and_step(scenes, True)
    assert rep_g["inner_iterations"] == rep_d["inner_iterations"] > 0
    assert np.array_equal(x_g, x_d)
    assert srep_g["inner_iterations"] == srep_d["inner_iterations"]
    for k in ("positions", "rotations", "velocities"):
        assert np.array_equal(out_g[k], out_d[k]), k


def test_direct_launch_rows_solve_equals_graph_replay(gpu, scenes):
    from vxsim import grid, solver
    s = scenes.large_64cube(dims=(16, 16, 12))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    rows = gs.export_rows()
    gs.close()
    inv = np.linalg.inv(rows["diagonal"].reshape(-1, 3, 3)).reshape(-1, 9)
    args = (np.ascontiguousarray(rows["rhs"], np.float32), np.ascontiguousarray(rows["neighbors"], np.uint32),
            np.ascontiguousarray(rows["diagonal"], np.float32), np.ascontiguousarray(rows["blocks"], np.float32),
            np.ascontiguousarray(inv, np.float32))
    ref = solver.solve_pcg(*args, 4000, 1e-5)
    grid.set_launch_mode(True)
    try:
        got = solver.solve_pcg(*args, 4000, 1e-5)
    finally:
        grid.set_launch_mode(False)
    assert got[1] == ref[1] and got[2] == ref[2]
    assert np.array_equal(got[0], ref[0])


def test_timed_solve_equals_solve_and_times_every_dispatch(gpu, scenes):
    s = scenes.large_64cube(dims=(24, 24, 24))
    cfg = s.config
    gs, _ = device_assembled(s, cfg.fixed_dt, cfg.gravity)
    rep = gs.solve(tolerance=1e-9, inner_tolerance=3e-3)
    x = gs.download()["correction"]
    timed = gs.solve_timed(tolerance=1e-9, inner_tolerance=3e-3)
    xt = gs.download()["correction"]
    again = gs.solve(tolerance=1e-9, inner_tolerance=3e-3)  # graphs still work after a timed solve
    gs.close()
    assert np.array_equal(x, xt)
    assert timed["inner_iterations"] == rep["inner_iterations"] == again["inner_iterations"]
    for k in ("spmv", "update"):
        t = timed[k]
        # every inner iteration's dispatch, plus those a chunk issues after the end
        assert t["dispatches"] >= rep["inner_iterations"], (k, t)
        assert 0 < t["min_us"] <= t["median_us"] <= t["max_us"] < 1e5
        assert t["min_us"] <= t["mean_us"] <= t["max_us"]
        assert t["total_us"] == pytest.approx(t["mean_us"] * t["dispatches"], rel=1e-9)
