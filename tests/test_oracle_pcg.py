"""PCG pieces of the oracle against the known answers of the reference's own
device tests and probe (src/cuda_solver.rs:507-554,
validation/cuda_sparse_probe.py) and against scipy on assembled systems."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EMPTY = 0xFFFFFFFF
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")


@pytest.fixture(scope="module")
def known():
    with open(GOLDEN) as f:
        return json.load(f)


def two_node_rows(known):
    g = known["two_node_pcg"]
    return dict(rhs=np.array(g["rhs"]).reshape(2, 3), neighbors=np.array(g["neighbors"], np.uint32).reshape(2, 6),
                diagonal=np.array(g["diagonal"]).reshape(2, 9), blocks=np.array(g["blocks"]).reshape(2, 54),
                inverse=np.array(g["inverse"]).reshape(2, 9))


def test_two_node_analytic_system(O, known):
    # validation/cuda_sparse_probe.py:268-343: x = ((2 b0 + 0.2 b1)/3.96, (0.2 b0 + 2 b1)/3.96).
    rows = two_node_rows(known)
    expect = np.array(known["two_node_pcg"]["expected"]).reshape(2, 3)
    for mode in (O.F32, O.F64):
        st, dx, it, rel = O.solve_device_semantics(rows, 80, 1e-5, mode)
        assert st == 0 and rel <= 1e-5
        assert np.abs(dx - expect).max() <= 1e-5
    x, _, rel = O.solve_exact(rows, 1e-14)
    np.testing.assert_allclose(x, expect, atol=1e-13)


def test_dot_and_ratio_known_answers(O, known):
    g = known["dot_600"]
    i = np.arange(3 * g["rows"])
    left = ((i % 97 + 1) * 0.03125).astype(np.float32)
    right = ((i % 53 + 1) * 0.015625).astype(np.float32)
    expected = float(np.dot(left.astype(np.float64), right.astype(np.float64)))
    assert expected == pytest.approx(g["expected"], rel=1e-15)
    dot = O.dot_rows(left, right)
    assert abs(dot - expected) <= max(abs(expected), 1.0) * 1e-12
    assert abs(O.guarded_ratio(expected * 0.75, dot) - 0.75) <= 1e-6
    r = known["ratio"]
    assert O.guarded_ratio(r["num"], r["den"]) == pytest.approx(r["expected"], abs=1e-6)
    assert O.guarded_ratio(153.0, 0.0) == 0.0
    assert O.guarded_ratio(1.0, 1e-31) == 0.0


def test_seventeen_partials_known_answer(O, known):
    """The reference's 17-partial reduction check (validation/cuda_sparse_probe.py:232-245):
    partials 1..17 reduce to 153. Here the partials are the per-256-row blocks
    of a row dot product (the first component of each block carries the value,
    the rest are zero), over 17 blocks as its 4,203-row captures use."""
    g = known["partial_sum_17"]
    vals = np.array(g["values"], dtype=np.float64)
    rows = 256 * len(vals)
    left = np.zeros(3 * rows)
    left[np.arange(len(vals)) * 3 * 256] = vals
    right = np.ones(3 * rows)
    assert O.dot_rows(left, right) == g["expected"] == float(vals.sum())
    # a duplicated or dropped partial would change the total
    assert O.dot_rows(left[: 3 * (rows - 256)], right[: 3 * (rows - 256)]) == g["expected"] - vals[-1]


def synthetic_rows(nx, ny, nz, known):
    g = known["synthetic_matvec"]
    n = nx * ny * nz
    nbr = np.full((n, 6), EMPTY, np.uint32)
    offsets = g["offsets"]
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = (z * ny + y) * nx + x
                for e, (dx, dy, dz) in enumerate(offsets):
                    a, b, c = x + dx, y + dy, z + dz
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        nbr[i, e] = (c * ny + b) * nx + a
    diag = np.tile(np.array(g["diagonal_block"], np.float64), (n, 1))
    blocks = np.tile(np.array(g["block"] * 6, np.float64), (n, 1))
    xv = (((np.arange(3 * n) % 19) - 9) * 0.03125).reshape(n, 3)
    return dict(neighbors=nbr, diagonal=diag, blocks=blocks), xv


def to_csr(rows):
    n = rows["diagonal"].shape[0]
    r, c, v = [], [], []
    for i in range(n):
        for a in range(3):
            for b in range(3):
                r.append(3 * i + a)
                c.append(3 * i + b)
                v.append(rows["diagonal"][i, 3 * a + b])
        for e in range(6):
            j = rows["neighbors"][i, e]
            if j == EMPTY:
                continue
            for a in range(3):
                for b in range(3):
                    r.append(3 * i + a)
                    c.append(3 * int(j) + b)
                    v.append(rows["blocks"][i, 9 * e + 3 * a + b])
    return sp.csr_matrix((v, (r, c)), shape=(3 * n, 3 * n))


def test_synthetic_matvec(O, known):
    # validation/cuda_sparse_probe.py:35-73 on a 20x25x20 lattice.
    rows, xv = synthetic_rows(20, 25, 20, known)
    A = to_csr(rows)
    expect = (A @ xv.ravel()).reshape(-1, 3)
    y64 = O.matvec(rows, xv, O.F64)
    np.testing.assert_allclose(y64, expect, rtol=1e-13, atol=1e-10)
    y32 = O.matvec(rows, xv, O.F32)
    assert np.abs(y32 - expect).max() <= 2e-3


@pytest.fixture(scope="module")
def c1_rows(O, scenes):
    s = scenes.tension_10cube()
    os_ = O.OracleScene(s)
    t = s.targets.astype(np.float64)
    m = ~np.isnan(t)
    os_.positions[m] = t[m]
    O.fit_rotations(os_, O.F64)
    sys_ = O.assemble(os_, O.F64, s.config.fixed_dt, os_.positions.copy(), s.free_mask(), None)
    return sys_, sys_.rows()


def test_exact_solve_matches_spsolve(O, c1_rows):
    _, rows = c1_rows
    A = to_csr(rows)
    b = rows["rhs"].ravel()
    x_ref = spla.spsolve(A.tocsc(), b)
    x, it, rel = O.solve_exact(rows, 1e-13)
    assert rel <= 1e-13
    assert np.linalg.norm(x.ravel() - x_ref) / np.linalg.norm(x_ref) <= 1e-10
    # The assembled matrix is symmetric positive definite.
    assert abs(A - A.T).max() <= 1e-9 * abs(A).max()


def test_cpu_pcg_preconditioners(O, c1_rows):
    sys_, rows = c1_rows
    x_ref, _, _ = O.solve_exact(rows, 1e-13)
    dx_bj, it_bj, res_bj, _ = sys_.solve_cpu(O.BJ, 2000, 1e-10)
    dx_ic, it_ic, res_ic, _ = sys_.solve_cpu(O.IC0, 2000, 1e-10)
    assert res_bj < 1e-10 and res_ic < 1e-10
    assert it_ic < it_bj  # IC(0) needs fewer iterations (src/solver.rs:1386-1389)
    for dx in (dx_bj, dx_ic):
        assert np.linalg.norm(dx - x_ref) / np.linalg.norm(x_ref) <= 1e-7


def test_device_semantics_check_interval_and_cap(O, c1_rows):
    _, rows = c1_rows
    st, dx, it, rel = O.solve_device_semantics(rows, 160, 1e-5)
    assert st == 0 and it % 16 == 0 and rel <= 1e-5
    # Checking every iteration stops earlier, at the first iteration below tol.
    st1, _, it1, rel1 = O.solve_device_semantics(rows, 160, 1e-5, check_every=1)
    assert st1 == 0 and it1 <= it and rel1 <= 1e-5
    st2, _, it2, rel2 = O.solve_device_semantics(rows, 7, 1e-9)
    assert st2 == 4 and it2 == 7 and rel2 > 1e-9  # not converged at the cap
    zero = dict(rows)
    zero["rhs"] = np.zeros_like(rows["rhs"])
    st3, dx3, it3, _ = O.solve_device_semantics(zero, 160, 1e-5)
    assert st3 == 0 and it3 == 0 and np.all(dx3 == 0)
    st4, _, _, _ = O.solve_device_semantics(rows, 0, 1e-5)
    assert st4 == 3  # the reference reports a non-finite residual with a zero cap
