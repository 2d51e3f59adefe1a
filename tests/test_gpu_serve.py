"""Serving loop (continuous batching): requests arriving while others run are
admitted as tiles free up, overlap on the device, and each result is the
same alone and in the mix (bit-identical to the scene stepped alone with the
same parameters, warm start included)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def alone(req):
    from vxsim import grid
    p = req.params
    with grid.GridScene(req.scene) as gs:
        for _ in range(req.steps):
            gs.step(fixed_dt=p.fixed_dt, gravity=tuple(p.gravity), substeps=p.substeps,
                    corotation_iterations=p.corotation_iterations, corotation_tolerance=p.corotation_tolerance,
                    solve=p.solve)
        return gs.download()


@pytest.fixture(scope="module")
def stream():
    from vxsim import serve
    # Arrivals every ~5 ms, boxes of 200..4000 REVs, 3 warm-started steps each.
    return serve.request_stream(rate=200.0, seconds=0.06, rev_lo=200, rev_hi=4000, steps=3, seed=11, max_requests=10)


def test_served_results_equal_results_alone(gpu, stream):
    from vxsim import batch, serve
    assert len(stream) >= 6
    with serve.Server(capacity_rows=64 * 512, max_scenes=8, keep_results=True) as S:
        S.run(stream)
    overlapped = 0
    for req in stream:
        assert req.state == batch.DONE, (req.rid, req.status, req.report)
        assert req.report["steps_done"] == req.steps
        ref = alone(req)
        for k in ("positions", "rotations", "velocities"):
            assert np.array_equal(req.result[k], ref[k]), (req.rid, k)
        overlapped += any(o is not req and o.admitted < req.finished and req.admitted < o.finished for o in stream)
    assert overlapped >= 2  # requests really shared the device


def test_admission_waits_for_free_tiles(gpu, stream):
    from vxsim import batch, serve
    biggest = max(-(-r.scene.n // batch.TILE) for r in stream)
    # Room for about two of the largest boxes: later requests must wait.
    with serve.Server(capacity_rows=2 * biggest * 512, max_scenes=4, policy="fifo") as S:
        S.run(stream)
    assert all(r.state == batch.DONE for r in stream)
    assert all(r.admitted >= r.arrival and r.finished > r.admitted for r in stream)
    waited = [r for r in stream if r.admitted - r.arrival > 1e-3]
    assert waited, "expected some requests to queue for tiles"
    p = serve.percentiles([r.latency for r in stream])
    assert 0 < p[50] <= p[99]


def test_served_red_black_results_equal_results_alone(gpu):
    """The same with the red-black preconditioner: warm starts kept on the
    device, requests overlapping, results bit-identical to each alone."""
    from vxsim import batch, grid, serve
    reqs = serve.request_stream(rate=200.0, seconds=0.06, rev_lo=200, rev_hi=4000, steps=3, seed=13, max_requests=8,
                                solve=grid.solve_params(tolerance=1e-6, warm_start=True, preconditioner=grid.RED_BLACK))
    with serve.Server(capacity_rows=64 * 512, max_scenes=8, keep_results=True) as S:
        S.run(reqs)
    for req in reqs:
        assert req.state == batch.DONE, (req.rid, req.status, req.report)
        ref = alone(req)
        for k in ("positions", "rotations", "velocities"):
            assert np.array_equal(req.result[k], ref[k]), (req.rid, k)


def test_sjf_holds_the_larger_scenes_and_results_equal_results_alone(gpu, stream):
    """Shortest job first with room for about two small scenes per tick: the
    larger scenes are held (no work, state kept in HBM) while smaller ones run,
    then resumed; every result is still bit-identical to the scene alone, and
    the smallest request is back before the largest."""
    from vxsim import batch, serve
    sizes = sorted(r.scene.n for r in stream)
    with serve.Server(capacity_rows=64 * 512, max_scenes=8, keep_results=True, policy="sjf",
                      active_revs=sizes[1] + sizes[2]) as S:
        S.run(stream)
    for req in stream:
        assert req.state == batch.DONE, (req.rid, req.status, req.report)
        ref = alone(req)
        for k in ("positions", "rotations", "velocities"):
            assert np.array_equal(req.result[k], ref[k]), (req.rid, k)
    small = min(stream, key=lambda r: r.scene.n)
    large = max(stream, key=lambda r: r.scene.n)
    assert small.finished < large.finished
    finishes = sorted(r.finished for r in stream)
    assert large.finished >= finishes[len(finishes) // 2]


def test_policy_arguments_are_checked(gpu):
    from vxsim import serve
    with pytest.raises(ValueError):
        serve.Server(capacity_rows=512, policy="lifo")
    with pytest.raises(ValueError):
        serve.Server(capacity_rows=512, active_revs=0)
