"""GPU parity of the streamed inner kernels. An arena whose tiles in use hold
more inner-plane bytes than the 256 MB Infinity Cache (here 128 x 128 x 96:
1.57 M rows, 3,072 tiles) runs the bond SpMV and the update with
nontemporal loads and stores; the arithmetic is the ordinary kernels'. The
solutions of block-Jacobi, red-black and multigrid solves satisfy the
oracle's FP64 system to the solve's tolerance."""
import numpy as np
import pytest

from test_gpu_grid import device_assembled, oracle_rows

pytestmark = pytest.mark.gpu


def test_streamed_solves_satisfy_the_oracle_system(gpu, O, scenes):
    from vxsim import grid
    s = scenes.large_64cube(dims=(128, 128, 96))
    assert -(-s.n // 512) >= 2560  # above the streaming threshold
    h, g = s.config.fixed_dt, s.config.gravity
    print("streamed solves: %s, %d rows" % (s.dims, s.n), flush=True)
    gs, _ = device_assembled(s, h, g)
    out = {}
    for name, pc in (("block-Jacobi", grid.BLOCK_JACOBI), ("red-black", grid.RED_BLACK),
                     ("multigrid", grid.MULTIGRID)):
        rep = gs.solve(grid.solve_params(tolerance=1e-9, max_inner=40000, preconditioner=pc))
        assert rep["status"] == 0 and rep["relative_residual"] <= 1e-9, (name, rep)
        out[name] = (rep, gs.download()["correction"])
    rows_g = gs.export_rows()
    gs.close()
    _, _, rows_o, _ = oracle_rows(O, s, h, g)
    assert np.array_equal(rows_g["active"], rows_o["active"])
    nb = np.linalg.norm(rows_o["rhs"])
    for name, (rep, x) in out.items():
        r = rows_o["rhs"] - O.matvec(rows_o, x[rows_g["active"]])
        rel = np.linalg.norm(r) / nb
        print("%s: %d inner, FP64 residual of the oracle system %.2e" % (name, rep["inner_iterations"], rel))
        assert rel <= 1e-9 * 1.01, name
