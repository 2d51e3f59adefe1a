"""Launch shapes of a one-scene arena (vx_scene_set_shape: the z-column
group of the tile order, the XCD deal of the inner kernels and the update
kernel's threads per tile) change which workgroup runs which tile and how
many rows each thread holds, never what a tile computes: a step under every
shape equals the default one bit for bit, on a block whose z-planes span
enough tiles for the grouping to act (18 tiles per plane), below and above
the 1,024-tile size where the large-scene kernels take over."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(0, 1, 512), (4, 2, 256), (16, 3, 512), (1, 0, 256), (64, 1, 256), (8, 0, 256)]


def _step(scenes, dims, shape):
    from vxsim import grid
    s = scenes.wood_block_blade(dims=dims)
    cfg = s.config
    sp = grid.solve_params(tolerance=1e-9, inner_tolerance=3e-3, max_inner=40000, max_outer=12)
    with grid.GridScene(s, tuned=False) as gs:
        if shape is not None:
            gs.set_shape(*shape)
        rep = gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                      corotation_iterations=2, corotation_tolerance=cfg.corotation_tolerance, solve=sp)
        return rep, gs.download()


@pytest.mark.parametrize("dims", [(96, 96, 8), (96, 96, 64)])
def test_every_shape_steps_bit_identically(gpu, scenes, dims):
    print("shape parity", dims, flush=True)
    rep0, out0 = _step(scenes, dims, None)
    assert rep0["inner_iterations"] > 0
    for shape in SHAPES:
        rep, out = _step(scenes, dims, shape)
        assert rep["inner_iterations"] == rep0["inner_iterations"], shape
        assert rep["outer_sweeps"] == rep0["outer_sweeps"], shape
        for k in ("positions", "velocities"):
            assert np.array_equal(out[k].view(np.uint64), out0[k].view(np.uint64)), (shape, k)


def test_shape_arguments_are_checked(gpu, scenes):
    from vxsim import grid
    from vxsim.solver import SolverError
    s = scenes.wood_block_blade(dims=(8, 8, 8))
    with grid.GridScene(s, tuned=False) as gs:
        for bad in ((3, 0, 512), (128, 0, 512), (8, 4, 512), (8, 0, 128), (8, 0, 1024)):
            with pytest.raises(SolverError):
                gs.set_shape(*bad)
        gs.set_shape(0, 3, 256)
