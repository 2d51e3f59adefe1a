"""GPU tests of the aggregation multigrid preconditioner
(VX_PRECOND_MULTIGRID) of the FP32 inner solve: solutions against the FP64
oracle's, iteration counts against block-Jacobi's on the same system,
repeated solves bit for bit, a step sequence against the oracle's FP64 steps,
and the argument checks."""
import numpy as np
import pytest

from test_gpu_grid import device_assembled, oracle_rows

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


SCENES = {
    "tension10": lambda sc: sc.tension_10cube(),
    "mixed20": lambda sc: sc.large_64cube(dims=(20, 18, 16)),
    "odd_box": lambda sc: sc.batch_boxes(count=3, seed=9)[2],
    "wood_blade": lambda sc: sc.wood_block_blade(dims=(48, 40, 24)),
    "thin_plate": lambda sc: sc.large_64cube(dims=(40, 36, 3)),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_multigrid_solve_matches_fp64_oracle(gpu, O, scenes, name):
    from vxsim import grid
    s = SCENES[name](scenes)
    h, g = s.config.fixed_dt, s.config.gravity
    print("multigrid solve: %s %s" % (name, s.dims), flush=True)
    gs, _ = device_assembled(s, h, g)
    bj = gs.solve(tolerance=1e-9, inner_tolerance=3e-3, max_inner=20000)
    mg = gs.solve(tolerance=1e-9, inner_tolerance=3e-3, max_inner=20000, preconditioner=grid.MULTIGRID)
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    assert np.array_equal(rows_g["active"], rows_o["active"])
    x_o, _, rel_o = O.solve_exact(rows_o, 1e-13)
    assert rel_o <= 1e-13
    err = rel_l2(x[rows_g["active"]], x_o)
    print("%s: block-Jacobi %d inner, multigrid %d inner + %d sweeps, residual %.2e, L2 error %.2e" %
          (name, bj["inner_iterations"], mg["inner_iterations"], mg["outer_sweeps"], mg["relative_residual"], err))
    assert mg["status"] == 0 and mg["relative_residual"] <= 1e-9
    assert err <= 1e-6
    assert mg["inner_iterations"] < bj["inner_iterations"]


def test_multigrid_64cube_first_solve(gpu, O, scenes):
    """The bench workload's first system: the error against the FP64 oracle
    and at least eight times fewer iterations than block-Jacobi."""
    from vxsim import grid
    s = scenes.large_64cube()
    h, g = s.config.fixed_dt, s.config.gravity
    gs, _ = device_assembled(s, h, g)
    bj = gs.solve(tolerance=1e-9, inner_tolerance=3e-3, max_inner=8000)
    mg = gs.solve(tolerance=1e-9, inner_tolerance=3e-3, max_inner=8000, preconditioner=grid.MULTIGRID)
    x = gs.download()["correction"]
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    x_o, _, _ = O.solve_exact(rows_o, 1e-13)
    err = rel_l2(x[rows_g["active"]], x_o)
    print("64^3: block-Jacobi %d inner, multigrid %d inner + %d sweeps, L2 error %.2e" %
          (bj["inner_iterations"], mg["inner_iterations"], mg["outer_sweeps"], err))
    assert mg["relative_residual"] <= 1e-9 and err <= 1e-6
    assert 8 * mg["inner_iterations"] <= bj["inner_iterations"]


def test_multigrid_solves_repeat_bit_for_bit(gpu, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(24, 20, 18))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    r1 = gs.solve(tolerance=1e-9, preconditioner=grid.MULTIGRID)
    x1 = gs.download()["correction"]
    r2 = gs.solve(tolerance=1e-9, preconditioner=grid.MULTIGRID)
    x2 = gs.download()["correction"]
    gs.close()
    assert r1["inner_iterations"] == r2["inner_iterations"] and np.array_equal(x1, x2)


def test_multigrid_steps_match_oracle_steps(gpu, O, scenes):
    """Three steps with two corotation passes each (the hierarchy is rebuilt
    after every assembly) against the oracle's FP64 steps."""
    from vxsim import grid
    s = scenes.large_64cube(dims=(12, 12, 12))
    cfg = s.config
    gs = grid.GridScene(s)
    sp = grid.solve_params(tolerance=1e-11, inner_tolerance=1e-3, preconditioner=grid.MULTIGRID)
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


def test_multigrid_arguments(gpu, scenes):
    from vxsim import batch, grid
    from vxsim.solver import SolverError
    s = scenes.large_64cube(dims=(8, 8, 8))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    with pytest.raises(SolverError):  # FP32 inner solve only
        gs.solve(tolerance=1e-9, mode=grid.F64, preconditioner=grid.MULTIGRID)
    with pytest.raises(SolverError):
        gs.solve(tolerance=1e-9, preconditioner=3)
    gs.close()
    c = s.config
    with batch.Batch(4 * batch.TILE, max_scenes=2) as B:  # one-scene arenas only
        with pytest.raises(SolverError):
            B.admit(s, 1, grid.step_params(fixed_dt=c.fixed_dt, gravity=c.gravity, substeps=1,
                                           corotation_iterations=1, corotation_tolerance=c.corotation_tolerance,
                                           solve=grid.solve_params(preconditioner=grid.MULTIGRID)))


def galerkin(A, keys, n3):
    """P^T A P for 2x2x2 aggregation of the nodes at integer `keys`
    (component by component), with the coarse keys; A is scipy sparse over
    3 * len(keys) unknowns."""
    import scipy.sparse as sp
    ck = keys // 2
    u, agg = np.unique(ck, axis=0, return_inverse=True)
    agg = agg.reshape(-1)
    n = len(keys)
    P = sp.csr_matrix((np.ones(3 * n), (np.arange(3 * n), np.repeat(3 * agg, 3) + np.tile(np.arange(3), n))),
                      shape=(3 * n, 3 * len(u)))
    return (P.T @ A @ P).tocsr(), u


def rows_matrix(rows):
    import scipy.sparse as sp
    na = len(rows["active"])
    nbr = rows["neighbors"].astype(np.int64)
    I, J, V = [], [], []
    bi = np.arange(na)
    for q in range(6):
        ok = nbr[:, q] != 0xFFFFFFFF
        blk = rows["blocks"][:, 9 * q:9 * q + 9].reshape(na, 3, 3)
        for a in range(3):
            for b in range(3):
                I.append(3 * bi[ok] + a)
                J.append(3 * nbr[ok, q] + b)
                V.append(blk[ok, a, b])
    D = rows["diagonal"].reshape(na, 3, 3)
    for a in range(3):
        for b in range(3):
            I.append(3 * bi + a)
            J.append(3 * bi + b)
            V.append(D[:, a, b])
    return sp.csr_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))), shape=(3 * na, 3 * na))


def check_level(lev, Ac, ck):
    """The device level's diagonal blocks and couplings against the Galerkin
    product on every node that has a child in the product."""
    nx, ny, nz = lev["dims"]
    sym = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    A = Ac.toarray() if Ac.shape[0] <= 6000 else None
    scale = float(np.abs(Ac.data).max())
    pos = {tuple(k): i for i, k in enumerate(ck)}
    worst = 0.0
    for i, k in enumerate(ck):
        I = int(k[0] + nx * (k[1] + ny * k[2]))
        blk = Ac[3 * i:3 * i + 3, 3 * i:3 * i + 3].toarray() if A is None else A[3 * i:3 * i + 3, 3 * i:3 * i + 3]
        dev = np.array([[lev["diagonal"][I, sym[r][c]] for c in range(3)] for r in range(3)], np.float64)
        worst = max(worst, np.abs(dev - blk).max() / scale)
        for a in range(3):
            kn = list(k)
            kn[a] += 1
            j = pos.get(tuple(kn))
            ref = np.zeros((3, 3)) if j is None else -(
                Ac[3 * i:3 * i + 3, 3 * j:3 * j + 3].toarray() if A is None else A[3 * i:3 * i + 3, 3 * j:3 * j + 3])
            worst = max(worst, np.abs(lev["couplings"][I, a].astype(np.float64) - ref).max() / scale)
        assert lev["masks"][I] == 7
    return worst


def test_multigrid_levels_are_galerkin_products(gpu, scenes):
    """Levels 1 and 2 of the device hierarchy against P^T A P computed here
    from the device's own exported FP64 system (2x2x2 aggregates of the
    active rows; the scene has no per-axis targets, so every active row is
    free in every component): diagonal blocks and couplings to 1e-6 of the
    largest entry (FP32 storage)."""
    from vxsim import grid
    s = scenes.large_64cube(dims=(20, 18, 15))
    gs, _ = device_assembled(s, s.config.fixed_dt, s.config.gravity)
    with pytest.raises(Exception):
        gs.multigrid_level(1)  # no hierarchy before a multigrid solve
    gs.solve(tolerance=1e-6, preconditioner=grid.MULTIGRID)
    rows = gs.export_rows()
    lev1, lev2 = gs.multigrid_level(1), gs.multigrid_level(2)
    past = gs.multigrid_level(20)
    gs.close()
    assert lev1["dims"] == (10, 9, 8) and lev2["dims"] == (5, 5, 4) and past is None
    keys = (np.asarray(s.keys) - np.asarray(s.keys).min(axis=0))[rows["active"]]  # lattice coordinates from 0
    A1, ck1 = galerkin(rows_matrix(rows), keys, 3)
    A2, ck2 = galerkin(A1, ck1, 3)
    w1, w2 = check_level(lev1, A1, ck1), check_level(lev2, A2, ck2)
    print("Galerkin levels: max deviation %.2e, %.2e of the largest entry" % (w1, w2))
    assert w1 <= 1e-6 and w2 <= 1e-6
