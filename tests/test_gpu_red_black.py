"""GPU parity of the red-black reduced-system preconditioner
(VX_PRECOND_RED_BLACK) of the structured solve, through the C ABI: its
solutions match the FP64 oracle's like block-Jacobi's do, its FP64-mode
iteration count is the oracle's red-black reduced CG count
(vxo_solve_red_black, itself pinned against dense red-black SGS-PCG in
test_oracle_red_black.py), it takes about half of block-Jacobi's
iterations, steps match the oracle's step, and red-black and block-Jacobi
scenes share a batch arena bit-identically to running alone."""
import numpy as np
import pytest

from test_gpu_grid import SCENES, device_assembled, oracle_rows

pytestmark = pytest.mark.gpu


def rb(grid, **kw):
    return grid.solve_params(preconditioner=grid.RED_BLACK, **kw)


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("mode", ["mixed", "f64"])
def test_red_black_solve_matches_fp64_oracle(gpu, O, scenes, name, mode):
    from vxsim import grid
    s = SCENES[name](scenes)
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    rep = gs.solve(rb(grid, tolerance=1e-10, inner_tolerance=1e-3, mode=grid.MIXED if mode == "mixed" else grid.F64))
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    x_o, _, _ = O.solve_exact(rows_o, 1e-13)
    xa = x[rows_g["active"]]
    assert rep["relative_residual"] <= 1e-10
    assert np.linalg.norm(xa - x_o) / np.linalg.norm(x_o) <= 1e-6
    inactive = np.setdiff1d(np.arange(s.n), rows_g["active"])
    assert np.all(x[inactive] == 0)
    # the full system's FP64 residual, red rows included, from the oracle's matvec
    r = rows_o["rhs"] - O.matvec(rows_o, xa)
    assert np.linalg.norm(r) / np.linalg.norm(rows_o["rhs"]) <= 2e-10


@pytest.mark.parametrize("name", list(SCENES))
def test_fp64_iterations_match_the_oracle_reduced_cg(gpu, O, scenes, name):
    """Plain FP64 mode runs one inner segment to half the tolerance; its count
    is the oracle's reduced CG count under the same stopping rule, and about
    half of block-Jacobi's."""
    from vxsim import grid
    s = SCENES[name](scenes)
    h, g = s.config.fixed_dt, s.config.gravity
    tol = 1e-9
    gs, _ = device_assembled(s, h, g)
    r_rb = gs.solve(rb(grid, tolerance=tol, mode=grid.F64))
    r_bj = gs.solve(tolerance=tol, mode=grid.F64)
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    black = O.black_rows(rows_g["active"], s.dims)
    _, k_o, rel_o, _ = O.solve_red_black(rows_o, black, tol, inner_tolerance=0.5 * tol)
    assert rel_o <= tol
    assert abs(r_rb["inner_iterations"] - k_o) <= 1, (r_rb["inner_iterations"], k_o)
    assert r_rb["inner_iterations"] <= 0.6 * r_bj["inner_iterations"]


def test_mixed_iterations_about_half_of_block_jacobi(gpu, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(32, 30, 28))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    r_bj = gs.solve(tolerance=1e-6)
    r_rb = gs.solve(rb(grid, tolerance=1e-6))
    gs.close()
    assert r_bj["relative_residual"] <= 1e-6 and r_rb["relative_residual"] <= 1e-6
    assert 0.35 * r_bj["inner_iterations"] <= r_rb["inner_iterations"] <= 0.6 * r_bj["inner_iterations"]


def test_red_black_step_matches_oracle_step(gpu, O, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(12, 12, 12))
    cfg = s.config
    gs = grid.GridScene(s)
    sp = rb(grid, tolerance=1e-11, inner_tolerance=1e-3)
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


def test_deterministic_and_warm_start(gpu, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(16, 16, 16))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    r1 = gs.solve(rb(grid, tolerance=1e-9))
    x1 = gs.download()["correction"]
    r2 = gs.solve(rb(grid, tolerance=1e-9))
    x2 = gs.download()["correction"]
    assert r1["inner_iterations"] == r2["inner_iterations"] and np.array_equal(x1, x2)
    r3 = gs.solve(rb(grid, tolerance=1e-9, warm_start=True))
    assert r3["inner_iterations"] == 0 and r3["outer_sweeps"] == 1
    gs.close()


def test_large_scene_pre_summed_reductions(gpu, O, scenes):
    """96 x 96 x 64 (1152 tiles): the pre-summed partials path with the
    red-black kernels; the correction satisfies the oracle's FP64 system."""
    from vxsim import grid
    s = scenes.large_64cube(dims=(96, 96, 64))
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    rep = gs.solve(rb(grid, tolerance=1e-9))
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    r = rows_o["rhs"] - O.matvec(rows_o, x[rows_g["active"]])
    assert np.linalg.norm(r) / np.linalg.norm(rows_o["rhs"]) <= 1e-9 * 1.01
    assert rep["relative_residual"] <= 1e-9 and rep["inner_iterations"] > 0


def test_invalid_preconditioner_is_rejected(gpu, scenes):
    from vxsim import grid
    from vxsim.solver import SolverError
    s = scenes.large_64cube(dims=(6, 5, 4))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    with pytest.raises(SolverError):  # 0, 1 and 2 (multigrid) are the preconditioners
        gs.solve(grid.solve_params(tolerance=1e-9, preconditioner=3))
    gs.close()


def test_batch_mixes_preconditioners(gpu, scenes):
    """Red-black and block-Jacobi scenes (mixed and FP64) in one arena end
    bit-identical to each stepped alone with its own preconditioner."""
    from vxsim import batch, grid
    from test_gpu_batch import alone, assert_same, params
    boxes = scenes.batch_boxes(count=5, seed=11, lo=5, hi=18)
    kinds = [dict(preconditioner=grid.RED_BLACK), dict(), dict(preconditioner=grid.RED_BLACK, mode=grid.F64),
             dict(mode=grid.F64), dict(preconditioner=grid.RED_BLACK)]
    total = sum(-(-s.n // batch.TILE) * batch.TILE for s in boxes)
    with batch.Batch(total, max_scenes=len(boxes)) as B:
        slots = [B.admit(s, 2, params(grid, s, **kw)) for s, kw in zip(boxes, kinds)]
        B.run()
        got = [(B.report(k), B.download(k)) for k in slots]
    for s, kw, (r, out) in zip(boxes, kinds, got):
        assert r["state"] == batch.DONE and r["last_relative_residual"] <= 1e-9, r
        ref, reps = alone(s, 2, tolerance=1e-9, **kw)
        assert r["inner_iterations"] == sum(x["inner_iterations"] for x in reps)
        assert_same(out, ref)


def test_batch_ticks_follow_the_iterations(gpu, scenes):
    """A red-black scene's inner chunks pipeline like block-Jacobi's (the next
    chunk is issued while the previous one runs), so a batch that needs half
    the iterations needs markedly fewer ticks."""
    from vxsim import batch, grid
    from test_gpu_batch import params
    boxes = scenes.batch_boxes(count=4, seed=5, lo=16, hi=24)
    total = sum(-(-s.n // batch.TILE) * batch.TILE for s in boxes)
    ticks, iters = {}, {}
    for name, pc in (("bj", grid.BLOCK_JACOBI), ("rb", grid.RED_BLACK)):
        with batch.Batch(total, max_scenes=len(boxes)) as B:
            slots = [B.admit(s, 1, params(grid, s, preconditioner=pc)) for s in boxes]
            ticks[name] = B.run()
            iters[name] = sum(B.report(k)["inner_iterations"] for k in slots)
    assert iters["rb"] <= 0.6 * iters["bj"]
    assert ticks["rb"] <= 0.8 * ticks["bj"], (ticks, iters)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (1, 1, 2), (3, 1, 1), (1, 2, 3), (2, 2, 2)])
def test_tiny_boxes(gpu, O, scenes, dims):
    """Boxes of one to eight REVs under random loads: an empty or one-row
    black system, red rows recovered exactly. Red-black and block-Jacobi
    solves both meet the oracle's FP64 solution."""
    from vxsim import grid
    s = scenes.uniform_box(dims, "iron")
    rng = np.random.default_rng(sum(dims))
    s.loads = (rng.standard_normal((s.n, 3)) * 0.1).astype(np.float32)
    h, g = 1e-3, (0.0, 0.0, 0.0)
    gs, _ = device_assembled(s, h, g)
    rows_g = gs.export_rows()
    out = {}
    for name, pc in (("bj", grid.BLOCK_JACOBI), ("rb", grid.RED_BLACK)):
        rep = gs.solve(grid.solve_params(tolerance=1e-10, preconditioner=pc))
        assert rep["relative_residual"] <= 1e-10, (name, rep)
        out[name] = gs.download()["correction"][rows_g["active"]]
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    x_o, _, _ = O.solve_exact(rows_o, 1e-13)
    for name, x in out.items():
        assert np.linalg.norm(x - x_o) <= 1e-6 * np.linalg.norm(x_o), name
