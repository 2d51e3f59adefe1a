"""The oracle's red-black reduced solve (vxo_solve_red_black), pinned by
what the mathematics fixes rather than by the reference, which has no
red-black preconditioner:

* its residual history equals that of CG preconditioned by red-black block
  symmetric Gauss-Seidel, M = (D + L) D^-1 (D + U) in red-first order, built
  densely here and started where the red residual is zero (Eisenstat's
  equivalence);
* its solution is numpy's dense solve of the assembled system;
* the red rows of its residual vanish (exact back-substitution);
* it needs about half the block-Jacobi iterations (the preconditioned
  spectrum of red-black SGS is the square of block-Jacobi's).

The systems are the oracle's FP64 assembly of mixed-material boxes."""
import numpy as np
import pytest


def assembled(O, scenes, dims):
    s = scenes.large_64cube(dims=dims)
    os_ = O.OracleScene(s)
    h = s.config.fixed_dt
    kin = s.kinematic != 0
    os_.positions[~kin] += np.array(s.config.gravity) * h * h
    O.fit_rotations(os_, O.F64)
    rows = O.assemble(os_, O.F64, h, os_.positions.copy(), s.free_mask(), s.loads).rows()
    return rows, O.black_rows(rows["active"], dims)


def dense(rows):
    n = rows["rhs"].shape[0]
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] = rows["diagonal"][i].reshape(3, 3)
        for e in range(6):
            j = int(rows["neighbors"][i, e])
            if j != 0xFFFFFFFF:
                A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = rows["blocks"][i, 9 * e:9 * e + 9].reshape(3, 3)
    return A, rows["rhs"].ravel().copy()


def test_colouring_splits_every_bond(O, scenes):
    rows, black = assembled(O, scenes, (5, 4, 3))
    for i in range(rows["rhs"].shape[0]):
        for j in rows["neighbors"][i]:
            if j != 0xFFFFFFFF:
                assert black[i] != black[j]


@pytest.mark.parametrize("dims", [(4, 3, 3), (6, 5, 4)])
def test_history_equals_red_black_sgs_pcg(O, scenes, dims):
    rows, black = assembled(O, scenes, dims)
    A, b = dense(rows)
    n = rows["rhs"].shape[0]
    red3 = np.repeat(black == 0, 3)
    order = np.concatenate([np.nonzero(red3)[0], np.nonzero(~red3)[0]])
    P = A[np.ix_(order, order)]
    # block diagonal D (3x3 blocks; the red-red and black-black parts of A are
    # block diagonal), strictly lower / upper parts in red-first order
    D = np.zeros_like(P)
    for i in range(n):
        D[3 * i:3 * i + 3, 3 * i:3 * i + 3] = P[3 * i:3 * i + 3, 3 * i:3 * i + 3]
    nr = int(red3.sum())
    assert np.abs(P[:nr, :nr] - D[:nr, :nr]).max() == 0 and np.abs(P[nr:, nr:] - D[nr:, nr:]).max() == 0
    L = np.tril(P - D)
    U = np.triu(P - D)
    M = (D + L) @ np.linalg.solve(D, D + U)
    bp = b[order]
    x = np.zeros_like(bp)
    x[:nr] = np.linalg.solve(D[:nr, :nr], bp[:nr])  # red residual zero at the start
    r = bp - P @ x
    hist = [np.linalg.norm(r)]
    z = np.linalg.solve(M, r)
    p = z.copy()
    rz = r @ z
    for _ in range(200):
        ap = P @ p
        a = rz / (p @ ap)
        x += a * p
        r -= a * ap
        hist.append(np.linalg.norm(r))
        if hist[-1] <= 1e-11 * np.linalg.norm(b):
            break
        z = np.linalg.solve(M, r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    _, k, rel, h_or = O.solve_red_black(rows, black, 1e-11, history=True)
    assert k == len(hist) - 1
    scale = hist[0]
    assert np.abs(h_or - np.array(hist)).max() <= 1e-8 * scale
    assert rel <= 1e-11


def test_solution_is_the_dense_solve_and_red_residual_vanishes(O, scenes):
    rows, black = assembled(O, scenes, (6, 5, 4))
    A, b = dense(rows)
    x, k, rel, _ = O.solve_red_black(rows, black, 1e-12)
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x.ravel() - xd) <= 1e-9 * np.linalg.norm(xd)
    r = (b - A @ x.ravel()).reshape(-1, 3)
    assert np.linalg.norm(r[black == 0]) <= 1e-12 * np.linalg.norm(b)
    assert rel == pytest.approx(np.linalg.norm(r) / np.linalg.norm(b), rel=1e-6, abs=1e-15)


def test_half_the_block_jacobi_iterations(O, scenes):
    rows, black = assembled(O, scenes, (16, 14, 12))
    _, k_rb, rel_rb, _ = O.solve_red_black(rows, black, 1e-9)
    _, k_bj, rel_bj = O.solve_exact(rows, 1e-9)
    assert rel_rb <= 1e-9 and rel_bj <= 1e-9
    assert 0.4 * k_bj <= k_rb <= 0.6 * k_bj


def test_inner_tolerance_stops_on_the_reduced_residual(O, scenes):
    rows, black = assembled(O, scenes, (6, 5, 4))
    _, k, _, h = O.solve_red_black(rows, black, 1e-14, inner_tolerance=1e-3, history=True)
    assert h[k] <= 1e-3 * h[0] < h[k - 1]
