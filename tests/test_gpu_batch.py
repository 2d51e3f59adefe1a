"""Batched scenes on one tile arena (vx_batch_*): every scene of a mixed batch
ends bit-identical to the same scene stepped alone (vx_scene_step, itself
checked against the oracle's step in test_gpu_grid.py), admission respects
the arena, failures stay per scene, and scenes admitted while others run
(continuous batching) are unaffected by their neighbours."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def params(grid, s, tol=1e-9, **kw):
    cfg = s.config
    return grid.step_params(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                            corotation_iterations=cfg.corotation_iterations,
                            corotation_tolerance=cfg.corotation_tolerance,
                            solve=grid.solve_params(tolerance=tol, **kw))


def alone(s, steps, **kw):
    from vxsim import grid
    cfg = s.config
    with grid.GridScene(s) as gs:
        reps = [gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                        corotation_iterations=cfg.corotation_iterations,
                        corotation_tolerance=cfg.corotation_tolerance, solve=grid.solve_params(**kw))
                for _ in range(steps)]
        out = gs.download()
    return out, reps


def assert_same(a, b):
    for k in ("positions", "rotations", "velocities"):
        assert np.array_equal(a[k], b[k]), k


@pytest.fixture(scope="module")
def boxes(scenes):
    return scenes.batch_boxes(count=12, seed=3, lo=5, hi=20)


def test_mixed_batch_matches_scenes_alone(gpu, boxes):
    from vxsim import batch, grid
    steps = 2
    total = sum(-(-s.n // batch.TILE) * batch.TILE for s in boxes)
    with batch.Batch(total, max_scenes=len(boxes)) as B:
        slots = [B.admit(s, steps, params(grid, s)) for s in boxes]
        assert None not in slots
        ticks = B.run()
        assert ticks > 0
        for s, slot in zip(boxes, slots):
            r = B.report(slot)
            assert r["state"] == batch.DONE and r["steps_done"] == steps, r
            ref, reps = alone(s, steps, tolerance=1e-9)
            assert r["solves"] == sum(x["solves"] for x in reps)
            assert r["inner_iterations"] == sum(x["inner_iterations"] for x in reps)
            assert r["last_relative_residual"] <= 1e-9
            assert_same(B.download(slot), ref)


def test_admission_and_release(gpu, boxes):
    from vxsim import batch, grid
    a, b = boxes[0], boxes[1]
    ta, tb = -(-a.n // batch.TILE), -(-b.n // batch.TILE)
    with batch.Batch(max(ta, tb) * batch.TILE, max_scenes=4, xcd_chunk_tiles=1) as B:
        largest, total = B.free_rows()
        assert largest == total and total >= max(ta, tb) * batch.TILE
        s0 = B.admit(a, 1, params(grid, a))
        assert s0 is not None
        if total - ta * batch.TILE < tb * batch.TILE:
            assert B.admit(b, 1, params(grid, b)) is None  # no room: nothing admitted
        B.run()
        assert B.report(s0)["state"] == batch.DONE
        B.release(s0)
        assert B.report(s0)["state"] == batch.FREE
        assert B.free_rows()[1] == total
        s1 = B.admit(b, 1, params(grid, b))
        assert s1 is not None
        B.run()
        ref, _ = alone(b, 1, tolerance=1e-9)
        assert_same(B.download(s1), ref)


def test_failed_scene_does_not_disturb_others(gpu, boxes):
    from vxsim import batch, grid
    good, bad = boxes[2], boxes[3]
    with batch.Batch(8 * batch.TILE * 8, max_scenes=4) as B:
        sg = B.admit(good, 1, params(grid, good))
        # An inner cap of 2 iterations and no refinement sweep cannot reach 1e-12.
        sb = B.admit(bad, 1, params(grid, bad, tol=1e-12, max_inner=2, max_outer=0))
        B.run()
        assert B.report(sb)["state"] == batch.FAILED and B.report(sb)["status"] == 4
        assert isinstance(B.error(sb), Exception)
        assert B.report(sg)["state"] == batch.DONE
        ref, _ = alone(good, 1, tolerance=1e-9)
        assert_same(B.download(sg), ref)


def test_continuous_admission_mid_run(gpu, boxes):
    """A scene admitted while another is mid-solve, and a third admitted into
    the tiles released by a finished one, all match their runs alone."""
    from vxsim import batch, grid
    a, b, c = boxes[4], boxes[5], boxes[6]
    with batch.Batch(4 * batch.TILE * 64, max_scenes=3) as B:
        sa = B.admit(a, 3, params(grid, a))
        for _ in range(3):
            B.tick()
        sb = B.admit(b, 1, params(grid, b))
        while B.report(sb)["state"] == batch.RUNNING:
            B.tick()
        out_b = B.download(sb)
        B.release(sb)
        sc = B.admit(c, 2, params(grid, c))
        B.run()
        assert_same(out_b, alone(b, 1, tolerance=1e-9)[0])
        assert_same(B.download(sa), alone(a, 3, tolerance=1e-9)[0])
        assert_same(B.download(sc), alone(c, 2, tolerance=1e-9)[0])


def test_fp64_mode_in_batch(gpu, boxes):
    from vxsim import batch, grid
    s = boxes[7]
    with batch.Batch(s.n, max_scenes=1) as B:
        slot = B.admit(s, 1, params(grid, s, tol=1e-9, mode=grid.F64))
        B.run()
        r = B.report(slot)
        assert r["state"] == batch.DONE and r["last_relative_residual"] <= 1e-9
        assert_same(B.download(slot), alone(s, 1, tolerance=1e-9, mode=grid.F64)[0])


def test_many_slot_arena_matches_scenes_alone(gpu, boxes):
    """An arena of more than 256 scene slots takes its phase bits through a
    device copy instead of the kernel argument; its scenes (admitted into
    high slots too) still end bit-identical to the scenes stepped alone."""
    from vxsim import batch, grid
    some = boxes[:3]
    total = sum(-(-s.n // batch.TILE) * batch.TILE for s in some)
    with batch.Batch(total, max_scenes=300) as B:
        slots = [B.admit(s, 1, params(grid, s)) for s in some]
        B.run()
        got = [B.download(k) for k in slots]
        assert all(B.report(k)["state"] == batch.DONE for k in slots)
    for s, out in zip(some, got):
        ref, _ = alone(s, 1, tolerance=1e-9)
        assert_same(out, ref)
