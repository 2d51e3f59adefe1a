"""vx_solve_pcg_noalloc (the drop-in solve with every device buffer in the
library's static scratch area) against vx_solve_pcg through the C ABI: the
same correction bit for bit, the same iteration count, relative residual and
status, on two calls in a row (the second smaller than the first) and on a
capped solve; sizes beyond the scratch area are refused before the device is
touched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rows(O, scene):
    """The reference row packing of a scene's first solve (the oracle's f32
    assembly: predict with gravity, fit, assemble)."""
    os_ = O.OracleScene(scene)
    h = scene.config.fixed_dt
    t = scene.targets.astype(np.float64)
    m = ~np.isnan(t)
    kin = scene.kinematic != 0
    v = os_.velocities + np.array(scene.config.gravity) * h
    v[m] = 0.0
    os_.positions[~kin] += v[~kin] * h
    os_.positions[m] = t[m]
    O.fit_rotations(os_, O.F32)
    rows = O.assemble(os_, O.F32, h, os_.positions.copy(), scene.free_mask(), scene.loads).rows()
    return (np.ascontiguousarray(rows["rhs"], np.float32), np.ascontiguousarray(rows["neighbors"], np.uint32),
            np.ascontiguousarray(rows["diagonal"], np.float32), np.ascontiguousarray(rows["blocks"], np.float32),
            np.ascontiguousarray(rows["inverse"], np.float32))


def _both(solver, args, cap, tol):
    out = []
    for fn in (solver.solve_pcg, solver.solve_pcg_noalloc):
        try:
            out.append((0,) + fn(*args, cap, tol))
        except solver.SolverError as e:
            out.append((e.status, None, None, None))
    return out


def test_noalloc_solve_equals_dropin_on_two_calls(gpu, O, scenes):
    from vxsim import solver
    big = _rows(O, scenes.large_64cube(dims=(28, 26, 24)))
    small = _rows(O, scenes.large_64cube(dims=(12, 10, 9)))
    assert big[0].shape[0] > small[0].shape[0] > 500
    for args in (big, small):
        (s0, x0, it0, rel0), (s1, x1, it1, rel1) = _both(solver, args, 4000, 1e-5)
        assert s0 == s1 == 0
        assert it0 == it1 > 0 and rel0 == rel1 <= 1e-5
        assert np.array_equal(x0.view(np.uint32), x1.view(np.uint32))
    # The allocation-free calls back to back (the second smaller), then the
    # drop-in: the scratch area's earlier contents change nothing.
    xa = solver.solve_pcg_noalloc(*big, 4000, 1e-5)
    xb = solver.solve_pcg_noalloc(*small, 4000, 1e-5)
    ya = solver.solve_pcg(*big, 4000, 1e-5)
    yb = solver.solve_pcg(*small, 4000, 1e-5)
    for x, y in ((xa, ya), (xb, yb)):
        assert x[1] == y[1] and x[2] == y[2]
        assert np.array_equal(x[0].view(np.uint32), y[0].view(np.uint32))


def test_noalloc_capped_solve_has_the_dropin_status(gpu, O, scenes):
    from vxsim import solver
    args = _rows(O, scenes.large_64cube(dims=(14, 12, 10)))
    (s0, *_), (s1, *_) = _both(solver, args, 20, 1e-12)
    assert s0 == s1 == 4  # VX_ERR_NOT_CONVERGED, src/cuda_solver.rs:487-492
    (s0, *_), (s1, *_) = _both(solver, args, 0, 1e-5)
    assert s0 == s1 == 3  # max_iterations 0: the reference's non-finite residual error


def test_noalloc_refuses_sizes_beyond_the_scratch_area(gpu):
    from vxsim import solver
    n = 200_000  # about 144 MB of buffers for a 64 MiB area
    rhs = np.ones((n, 3), np.float32)
    nbr = np.full((n, 6), 0xFFFFFFFF, np.uint32)
    diag = np.tile(np.eye(3, dtype=np.float32).ravel(), (n, 1))
    blk = np.zeros((n, 54), np.float32)
    with pytest.raises(solver.SolverError) as e:
        solver.solve_pcg_noalloc(rhs, nbr, diag, blk, diag, 100, 1e-5)
    assert e.value.status == 2 and "scratch" in str(e.value)
    # The drop-in itself takes the same system (grow-only buffers).
    x, it, rel = solver.solve_pcg(rhs, nbr, diag, blk, diag, 100, 1e-5)
    assert np.allclose(x, 1.0)
