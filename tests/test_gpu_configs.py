"""The benchmark workloads themselves against the FP64 oracle, at their real
sizes, through the C ABI:

* the 64^3 mixed box: the first solve's relative L2 error against the
  oracle's FP64 solution of the same system (bar 1e-6);
* 256 independent boxes with edges U[8, 48] in one arena, one step each:
  every scene's correction satisfies the oracle-assembled FP64 system, and for
  eight scenes (the smallest, the largest and six spread over the batch) the
  correction and the configured two-pass step match the oracle's FP64
  solution / step within 1e-6 relative L2;
* the first twelve requests of the served stream (Poisson arrivals at 200/s,
  REVs log-uniform in [1e3, 2e6], aspect ratios U[0.5, 2], ten warm-started
  steps, admission against a 200 GB HBM budget): every result bit-identical to
  the same request stepped alone, and the three smallest requests' ten steps
  within 1e-6 relative L2 of the oracle's FP64 steps.

A relative residual of 1e-6 leaves up to 1e-3 relative L2 error on these
boxes (their condition numbers reach ~1e6), so the solves here and in the
benchmark's batched and served legs run to a relative residual of 1e-10, which
the error bars below check."""
import numpy as np
import pytest

from test_gpu_grid import device_assembled, oracle_rows

pytestmark = pytest.mark.gpu

BATCH_TOL = 1e-10


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_64cube_first_solve_error_vs_fp64_oracle(gpu, O, scenes):
    """The bench workload's first solve (relative residual 1e-9, FP32 inner
    segments of 3e-3) against the oracle's FP64 solution to 1e-13."""
    s = scenes.large_64cube()
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    rep = gs.solve(tolerance=1e-9, inner_tolerance=3e-3, max_inner=8000)
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    assert np.array_equal(rows_g["active"], rows_o["active"])
    x_o, _, rel_o = O.solve_exact(rows_o, 1e-13)
    assert rel_o <= 1e-13
    err = rel_l2(x[rows_g["active"]], x_o)
    print("64^3 first solve: %d inner + %d sweeps, residual %.3e, L2 error vs FP64 oracle %.3e" %
          (rep["inner_iterations"], rep["outer_sweeps"], rep["relative_residual"], err))
    assert rep["relative_residual"] <= 1e-9
    assert err <= 1e-6


def _batch_run(boxes, passes, tol):
    from vxsim import batch, grid
    total = sum(-(-s.n // batch.TILE) * batch.TILE for s in boxes)
    with batch.Batch(total, max_scenes=len(boxes)) as B:
        slots = []
        for s in boxes:
            c = s.config
            sp = grid.step_params(fixed_dt=c.fixed_dt, gravity=c.gravity, substeps=c.substeps,
                                  corotation_iterations=passes, corotation_tolerance=c.corotation_tolerance,
                                  solve=grid.solve_params(tolerance=tol, max_inner=20000))
            slots.append(B.admit(s, 1, sp))
        assert None not in slots
        B.run()
        reps = [B.report(k) for k in slots]
        outs = [B.download(k) for k in slots]
    return reps, outs


@pytest.fixture(scope="module")
def c3_boxes(scenes):
    boxes = scenes.batch_boxes(count=256, seed=3, lo=8, hi=48)
    ns = np.array([s.n for s in boxes])
    picked = sorted({int(ns.argmin()), int(ns.argmax())} | set(range(16, 256, 40)))
    return boxes, picked


def test_c3_256_boxes_single_pass_vs_fp64_oracle(gpu, O, c3_boxes):
    """One pass per box (predict, fit, assemble, solve, apply): the device
    correction is the step's position change, which must satisfy every box's
    oracle-assembled FP64 system to the requested residual, and match the
    oracle's FP64 solution on the picked boxes."""
    from vxsim import batch
    boxes, picked = c3_boxes
    assert len(boxes) == 256
    reps, outs = _batch_run(boxes, 1, BATCH_TOL)
    worst_res, errs = 0.0, {}
    for b, (s, r, out) in enumerate(zip(boxes, reps, outs)):
        assert r["state"] == batch.DONE and r["steps_done"] == 1, (b, r)
        assert r["last_relative_residual"] <= BATCH_TOL, (b, r)
        os_, _, rows, _ = oracle_rows(O, s, s.config.fixed_dt, s.config.gravity)
        act = rows["active"]
        dx = out["positions"][act] - os_.positions[act]  # os_.positions: the oracle's predicted positions
        res = np.linalg.norm(rows["rhs"] - O.matvec(rows, dx)) / np.linalg.norm(rows["rhs"])
        worst_res = max(worst_res, res)
        assert res <= 2 * BATCH_TOL, (b, s.dims, res)
        if b in picked:
            x_o, _, rel_o = O.solve_exact(rows, 1e-13)
            errs[b] = rel_l2(dx, x_o)
    print("C3 single pass: worst oracle residual %.3e; L2 errors %s" %
          (worst_res, {b: "%.2e" % e for b, e in errs.items()}))
    assert max(errs.values()) <= 1e-6


def test_c3_256_boxes_configured_step_vs_fp64_oracle(gpu, O, c3_boxes):
    """The configured step (up to two corotation passes with early exit) of
    all 256 boxes in one arena; the picked boxes' positions and velocities
    against the oracle's FP64 step with exact solves, and bit-identical to the
    box stepped alone."""
    from vxsim import batch, grid
    boxes, picked = c3_boxes
    reps, outs = _batch_run(boxes, 2, BATCH_TOL)
    assert all(r["state"] == batch.DONE for r in reps)
    errs = {}
    for b in picked:
        s, out = boxes[b], outs[b]
        cfg = s.config
        with grid.GridScene(s) as gs:
            gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps, corotation_iterations=2,
                    corotation_tolerance=cfg.corotation_tolerance,
                    solve=grid.solve_params(tolerance=BATCH_TOL, max_inner=20000))
            alone = gs.download()
        for k in ("positions", "rotations", "velocities"):
            assert np.array_equal(out[k], alone[k]), (b, k)
        os_ = O.OracleScene(s)
        start = os_.positions.copy()
        O.step(os_, cfg, O.F64, O.BJ, exact=True, exact_tolerance=1e-13, loads=s.loads)
        errs[b] = rel_l2(out["positions"] - start, os_.positions - start)
        assert errs[b] <= 1e-6, (b, s.dims, errs[b])
        np.testing.assert_allclose(out["velocities"], os_.velocities, rtol=0,
                                   atol=1e-6 * np.abs(os_.velocities).max())
    print("C3 configured step: displacement L2 errors %s" % {b: "%.2e" % e for b, e in errs.items()})


def test_c4_stream_in_the_mix_equals_alone_and_oracle(gpu, O):
    """The first twelve arrivals of the served stream at its size mix (0.98 K
    to 1.7 M REVs, 6.6 M in all), ten warm-started steps each, served with
    continuous admission in an arena sized from a 200 GB budget."""
    from vxsim import batch, grid, serve
    sp = grid.solve_params(tolerance=BATCH_TOL, warm_start=True, max_inner=40000)
    reqs = serve.request_stream(200.0, 60.0, 1e3, 2e6, 10, seed=4, max_requests=12, solve=sp)
    assert len(reqs) == 12
    with serve.Server.for_budget(200 << 30, max_scenes=256, keep_results=True) as S:
        S.run(reqs)
    overlapped = 0
    for req in reqs:
        assert req.state == batch.DONE, (req.rid, req.status, req.report)
        assert req.report["steps_done"] == 10 and req.report["last_relative_residual"] <= BATCH_TOL
        overlapped += any(o is not req and o.admitted < req.finished and req.admitted < o.finished for o in reqs)
    assert overlapped >= 2  # the requests shared the device
    for req in reqs:
        p = req.params
        with grid.GridScene(req.scene) as gs:
            for _ in range(req.steps):
                gs.step(fixed_dt=p.fixed_dt, gravity=tuple(p.gravity), substeps=p.substeps,
                        corotation_iterations=p.corotation_iterations, corotation_tolerance=p.corotation_tolerance,
                        solve=p.solve)
            alone = gs.download()
        for k in ("positions", "rotations", "velocities"):
            assert np.array_equal(req.result[k], alone[k]), (req.rid, k)
    errs = {}
    for req in sorted(reqs, key=lambda r: r.scene.n)[:3]:
        s = req.scene
        os_ = O.OracleScene(s)
        start = os_.positions.copy()
        cfg = s.config
        cfg.corotation_iterations = req.params.corotation_iterations
        cfg.corotation_tolerance = req.params.corotation_tolerance
        cfg.gravity = tuple(req.params.gravity)
        for _ in range(req.steps):
            O.step(os_, cfg, O.F64, O.BJ, exact=True, exact_tolerance=1e-13, loads=s.loads)
        errs[req.rid] = rel_l2(req.result["positions"] - start, os_.positions - start)
        assert errs[req.rid] <= 1e-6, (req.rid, s.dims, errs[req.rid])
    print("C4 stream: REVs %s; oracle L2 errors after ten steps %s" %
          ([r.scene.n for r in reqs], {k: "%.2e" % e for k, e in errs.items()}))
