"""GPU parity of the drop-in vx_solve_pcg / vx_solve_pcg_f64 against the
oracle's restatement of the reference device solve (src/cuda_solver.rs:
301-501, src/cuda/pcg.ptx), through the C ABI."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EMPTY = 0xFFFFFFFF


def rows_for(O, s, mode):
    os_ = O.OracleScene(s)
    h = s.config.fixed_dt / s.config.substeps
    t = s.targets.astype(np.float64)
    m = ~np.isnan(t)
    os_.positions[m] = t[m]
    kin = s.kinematic != 0
    os_.positions[~kin & ~m.any(axis=1)] += np.array(s.config.gravity) * h * h
    O.fit_rotations(os_, mode)
    return O.assemble(os_, mode, h, os_.positions.copy(), s.free_mask(), s.loads).rows()


def call(solver, rows, cap, tol, f64=False):
    fn = solver.solve_pcg_f64 if f64 else solver.solve_pcg
    return fn(rows["rhs"], rows["neighbors"], rows["diagonal"], rows["blocks"], rows["inverse"], cap, tol)


@pytest.fixture(scope="module")
def c1(O, scenes):
    return rows_for(O, scenes.tension_10cube(), O.F32)


@pytest.fixture(scope="module")
def mixed24(O, scenes):
    s = scenes.large_64cube(dims=(24, 24, 24))
    return rows_for(O, s, O.F32)


@pytest.mark.parametrize("name", ["c1", "mixed24"])
@pytest.mark.parametrize("tol", [1e-5, 1e-6])
def test_f32_matches_reference_device_semantics(gpu, O, request, name, tol):
    from vxsim import solver
    rows = request.getfixturevalue(name)
    st, dx_o, it_o, rel_o = O.solve_device_semantics(rows, 4000, tol)
    dx, it, rel = call(solver, rows, 4000, tol)
    assert st == 0
    assert it == it_o and it % 16 == 0
    assert rel == pytest.approx(rel_o, rel=1e-5)
    assert np.linalg.norm(dx - dx_o) / np.linalg.norm(dx_o) <= 1e-6


def test_f64_matches_fp64_solution(gpu, O, scenes):
    from vxsim import solver
    rows = rows_for(O, scenes.large_64cube(dims=(16, 16, 16)), O.F64)
    # The boundary takes f32 arrays: the FP64 reference solution is that of
    # the f32-rounded system, solved in double.
    f32rows = {k: (v.astype(np.float32).astype(np.float64) if v.dtype == np.float64 else v) for k, v in rows.items()}
    x_o, _, _ = O.solve_exact(f32rows, 1e-13)
    dx, it, rel = call(solver, f32rows, 20000, 1e-10, f64=True)
    assert rel <= 1e-10
    assert np.linalg.norm(dx - x_o) / np.linalg.norm(x_o) <= 1e-6


def test_two_node_known_answer(gpu):
    from vxsim import solver
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))["two_node_pcg"]
    rows = dict(rhs=np.array(g["rhs"]).reshape(2, 3), neighbors=np.array(g["neighbors"], np.uint32).reshape(2, 6),
                diagonal=np.array(g["diagonal"]).reshape(2, 9), blocks=np.array(g["blocks"]).reshape(2, 54),
                inverse=np.array(g["inverse"]).reshape(2, 9))
    for f64 in (False, True):
        dx, it, rel = call(solver, rows, 80, 1e-5, f64)
        assert np.abs(dx - np.array(g["expected"]).reshape(2, 3)).max() <= 1e-5


def test_not_converged_and_cap_semantics(gpu, O, c1):
    from vxsim import solver
    st_o, _, it_o, rel_o = O.solve_device_semantics(c1, 7, 1e-9)
    with pytest.raises(solver.SolverError) as e:
        call(solver, c1, 7, 1e-9)
    assert e.value.status == 4 and "did not converge" in str(e.value) and st_o == 4
    with pytest.raises(solver.SolverError) as e:
        call(solver, c1, 0, 1e-5)
    assert e.value.status == 3  # zero cap: the reference's non-finite residual error


def test_nonfinite_and_invalid_inputs(gpu, c1):
    from vxsim import solver
    bad = dict(c1)
    bad["rhs"] = c1["rhs"].copy()
    bad["rhs"][3, 1] = np.nan
    with pytest.raises(solver.SolverError) as e:
        call(solver, bad, 160, 1e-5)
    assert e.value.status == 3
    bad = dict(c1)
    bad["neighbors"] = c1["neighbors"].copy()
    bad["neighbors"][5, 0] = c1["neighbors"].shape[0] + 10
    with pytest.raises(solver.SolverError) as e:
        call(solver, bad, 160, 1e-5)
    assert e.value.status == 6
    # The library stays usable after errors.
    dx, it, rel = call(solver, c1, 160, 1e-5)
    assert rel <= 1e-5


def test_repeated_and_resized_calls_are_reproducible(gpu, c1, mixed24):
    from vxsim import solver
    a = call(solver, c1, 4000, 1e-6)
    b = call(solver, mixed24, 4000, 1e-6)
    c = call(solver, c1, 4000, 1e-6)
    d = call(solver, mixed24, 4000, 1e-6)
    assert np.array_equal(a[0], c[0]) and a[1] == c[1]
    assert np.array_equal(b[0], d[0]) and b[1] == d[1]


def rows_matvec(rows, x):
    """y = A x over the reference row packing (diagonal block + six slots), FP64."""
    n = rows["rhs"].shape[0]
    x = np.asarray(x, np.float64).reshape(n, 3)
    y = np.einsum("nrc,nc->nr", rows["diagonal"].reshape(n, 3, 3).astype(np.float64), x)
    nb = rows["neighbors"].reshape(n, 6)
    blk = rows["blocks"].reshape(n, 6, 3, 3).astype(np.float64)
    for s in range(6):
        j = nb[:, s]
        ok = j != EMPTY
        y[ok] += np.einsum("nrc,nc->nr", blk[ok, s], x[j[ok]])
    return y


def test_large_grid_pre_summed_partials(gpu, O, scenes):
    """More than 1,024 workgroups (580,608 rows): the partials of each kernel
    are pre-summed once instead of re-summed by every workgroup. The solve
    stops on a 16-iteration boundary with the residual it reports, that
    residual is the FP64 true residual of the correction it returns, and the
    correction agrees with the FP64 drop-in's solution of the same rows."""
    from vxsim import solver
    rows = rows_for(O, scenes.large_64cube(dims=(96, 96, 64)), O.F32)
    f32rows = {k: (v.astype(np.float32) if v.dtype == np.float64 else v) for k, v in rows.items()}
    n = f32rows["rhs"].shape[0]
    assert -(-n // 512) > 1024
    tol = 1e-4
    dx, it, rel = call(solver, f32rows, 4000, tol)
    assert it % 16 == 0 and rel <= tol
    b = f32rows["rhs"].astype(np.float64)
    true_rel = np.linalg.norm(b - rows_matvec(f32rows, dx)) / np.linalg.norm(b)
    assert true_rel == pytest.approx(rel, rel=0.1)
    x64, it64, rel64 = call(solver, f32rows, 20000, 1e-9, f64=True)
    assert rel64 <= 1e-9
    assert np.linalg.norm(dx - x64) / np.linalg.norm(x64) <= 50 * tol
