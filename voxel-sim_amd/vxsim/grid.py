"""Device-resident structured scenes (vx_scene_* of include/vx_solver.h).

A `GridScene` uploads a `vxsim.scenes.Scene` box once and then drives the
solve part of a physics step on the device: predict, rotation fit
(src/solver.rs:327-417), bond assembly (src/solver.rs:1542-1636) and the
mixed-precision PCG (FP32 inner block-Jacobi PCG, FP64 refinement).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib
from ._abi import (FractureParams, KernelTimes, SceneDesc, SolveParams, SolveReport, StepParams, StepReport,
                   VX_DEAL_AUTO, VX_DEAL_CONTIGUOUS, VX_DEAL_ROUND_ROBIN, VX_DEAL_SPLIT,
                   VX_LAUNCH_DIRECT, VX_LAUNCH_GRAPHS, VX_PRECOND_BLOCK_JACOBI, VX_PRECOND_MULTIGRID,
                   VX_PRECOND_RED_BLACK, VX_SOLVE_F64, VX_SOLVE_MIXED, c_dp, c_fp, c_u8p, c_u32p)
from .solver import SolverError, check

MIXED = VX_SOLVE_MIXED
F64 = VX_SOLVE_F64
BLOCK_JACOBI = VX_PRECOND_BLOCK_JACOBI
RED_BLACK = VX_PRECOND_RED_BLACK
MULTIGRID = VX_PRECOND_MULTIGRID


def solve_params(tolerance=1e-9, inner_tolerance=3e-3, max_inner=4000, max_outer=12, warm_start=False, mode=MIXED,
                 preconditioner=BLOCK_JACOBI):
    """vx_solve_params. `preconditioner`: BLOCK_JACOBI (the reference's),
    RED_BLACK (CG on the red-black reduced system: red-black block symmetric
    Gauss-Seidel, about half the iterations) or MULTIGRID (one aggregation
    multigrid V-cycle per iteration; one-scene arenas, mode MIXED only)."""
    return SolveParams(float(tolerance), float(inner_tolerance), int(max_inner), int(max_outer), 1 if warm_start else 0,
                       int(mode), int(preconditioner), 0)


def _report(r) -> dict:
    out = {}
    for f in r._fields_:
        v = getattr(r, f[0])
        out[f[0]] = _report(v) if isinstance(v, ctypes.Structure) else v
    return out


def fracture_params(config=None) -> FractureParams:
    """vx_fracture_params from a `vxsim.scenes.FractureConfig` (None: the
    pass is off)."""
    if config is None:
        return FractureParams()
    c = config
    return FractureParams(1, 0, float(c.damage_rate), float(c.energy_release), float(c.fatigue_cycles),
                          float(c.fatigue_exponent), float(c.fatigue_threshold), float(c.plastic_heat_fraction))


def scene_desc(scene):
    """A vx_scene_desc over the scene's arrays; returns (desc, arrays to keep
    alive while the description is in use)."""
    nx, ny, nz = scene.dims
    if nx * ny * nz != scene.n:
        raise ValueError("scene is not a full box")
    keep = []

    def f(a):
        if a is None:
            return None
        arr = np.ascontiguousarray(a, dtype=np.float32)
        keep.append(arr)
        return arr.ctypes.data_as(c_fp)

    kin = np.ascontiguousarray(scene.kinematic, np.uint8)
    nbr = np.ascontiguousarray(scene.neighbor, np.uint32)
    keep += [kin, nbr]
    desc = SceneDesc(
        nx, ny, nz, f(scene.positions), f(scene.rotations), f(scene.velocities), f(scene.mass),
        kin.ctypes.data_as(c_u8p), f(scene.young), f(scene.young_perp), f(scene.shear), f(scene.solid),
        f(scene.fibers), f(scene.anisotropy), f(scene.delta_temp), f(scene.alpha_par), f(scene.alpha_perp),
        nbr.ctypes.data_as(c_u32p), f(scene.damage), f(scene.kscale), f(scene.formation_offset),
        f(scene.plastic_strain), f(scene.targets), f(scene.loads), f(getattr(scene, "strength", None)),
        f(getattr(scene, "bond_history", None)))
    return desc, keep


def step_params(fixed_dt=0.001, gravity=(0.0, 0.0, -9.81), substeps=1, corotation_iterations=2,
                corotation_tolerance=1e-5, solve=None, fracture=None) -> StepParams:
    """vx_step_params; `fracture` is a `vxsim.scenes.FractureConfig` that
    turns the fracture pass on (None: off)."""
    return StepParams(float(fixed_dt), (ctypes.c_double * 3)(*gravity), int(substeps), int(corotation_iterations),
                      float(corotation_tolerance), solve if solve is not None else solve_params(),
                      fracture_params(fracture))


def download_bond_forces(fn, *args, n: int) -> dict:
    """Bond forces of the last substep through vx_scene_download_bond_forces /
    vx_batch_download_bond_forces: last_force and lambda, (3, n, 3) f64."""
    lf = np.empty((3, n, 3))
    lam = np.empty((3, n, 3))
    check(fn(*args, lf.ctypes.data_as(c_dp), lam.ctypes.data_as(c_dp)))
    return dict(last_force=lf, **{"lambda": lam})


def download_bonds(fn, *args, n: int) -> dict:
    """Bond state through vx_scene_download_bonds / vx_batch_download_bonds:
    neighbor (3, n), damage (3, n), plastic_strain (3, n, 3), history (3, n, 6)."""
    nb = np.empty((3, n), np.uint32)
    dmg = np.empty((3, n), np.float32)
    pl = np.empty((3, n, 3), np.float32)
    hist = np.empty((3, n, 6), np.float32)
    check(fn(*args, nb.ctypes.data_as(c_u32p), dmg.ctypes.data_as(c_fp), pl.ctypes.data_as(c_fp),
             hist.ctypes.data_as(c_fp)))
    return dict(neighbor=nb, damage=dmg, plastic_strain=pl, history=hist)


class GridScene:
    def __init__(self, scene, tuned: bool = True):
        """Upload `scene`; with `tuned`, the iterations per tick and the
        launch shape (z-column group, XCD deal) come from the schedule table
        for the scene's size (vxsim.tune)."""
        desc, keep = scene_desc(scene)
        nx, ny, nz = scene.dims
        handle = ctypes.c_void_p()
        check(lib().vx_scene_create(ctypes.byref(desc), ctypes.byref(handle)))
        self.handle = handle
        self.n = scene.n
        self.dims = (nx, ny, nz)
        self.active = scene.active_count()
        self._kinematic = np.asarray(scene.kinematic, np.uint8).copy()
        if tuned:
            from . import tune
            t = tune.lookup(scene.n, kind="single")
            self.set_schedule(t["iters"])
            self.set_shape(t["zcols"], t["deal"], t["update_threads"])

    def close(self):
        if getattr(self, "handle", None):
            lib().vx_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stages ---------------------------------------------------------------
    def predict(self, h, gravity=(0.0, 0.0, 0.0)):
        g = (ctypes.c_double * 3)(*gravity)
        check(lib().vx_scene_predict(self.handle, float(h), g))

    def fit_rotations(self) -> float:
        a = ctypes.c_double(0.0)
        check(lib().vx_scene_fit_rotations(self.handle, ctypes.byref(a)))
        return float(a.value)

    def assemble(self, h):
        check(lib().vx_scene_assemble(self.handle, float(h)))

    def solve(self, params=None, raise_on_error=True, **kw) -> dict:
        p = params if params is not None else solve_params(**kw)
        rep = SolveReport()
        status = lib().vx_scene_solve(self.handle, ctypes.byref(p), ctypes.byref(rep))
        if status and raise_on_error:
            check(status)
        out = _report(rep)
        out["status"] = status
        return out

    def apply(self):
        check(lib().vx_scene_apply(self.handle))

    def finish_substep(self, h):
        check(lib().vx_scene_finish_substep(self.handle, float(h)))

    def step(self, fixed_dt=0.001, gravity=(0.0, 0.0, -9.81), substeps=1, corotation_iterations=2,
             corotation_tolerance=1e-5, solve=None, fracture=None) -> dict:
        return self.step_with(step_params(fixed_dt, gravity, substeps, corotation_iterations, corotation_tolerance,
                                          solve, fracture))

    def step_with(self, params: StepParams) -> dict:
        """One fixed step with a prepared vx_step_params."""
        rep = StepReport()
        check(lib().vx_scene_step(self.handle, ctypes.byref(params), ctypes.byref(rep)))
        return _report(rep)

    # -- data -----------------------------------------------------------------
    def download(self) -> dict:
        n = self.n
        pos = np.empty((n, 3))
        rot = np.empty((n, 4))
        vel = np.empty((n, 3))
        x = np.empty((n, 3))
        check(lib().vx_scene_download(self.handle, pos.ctypes.data_as(c_dp), rot.ctypes.data_as(c_dp),
                                      vel.ctypes.data_as(c_dp), x.ctypes.data_as(c_dp)))
        return dict(positions=pos, rotations=rot, velocities=vel, correction=x)

    def set_targets_and_loads(self, targets=None, loads=None):
        """Axis targets (n, 3), NaN = free, and applied loads (n, 3) for the
        following steps (None: no targets / zero loads). Drops the warm start."""
        t = None if targets is None else np.ascontiguousarray(targets, np.float32).reshape(self.n, 3)
        l_ = None if loads is None else np.ascontiguousarray(loads, np.float32).reshape(self.n, 3)
        check(lib().vx_scene_set_targets_and_loads(self.handle, None if t is None else t.ctypes.data_as(c_fp),
                                                   None if l_ is None else l_.ctypes.data_as(c_fp)))
        free = (np.isnan(t) if t is not None else np.ones((self.n, 3), bool)) & (self._kinematic == 0)[:, None]
        self.active = int(np.count_nonzero(free.any(axis=1)))

    def download_bonds(self) -> dict:
        return download_bonds(lib().vx_scene_download_bonds, self.handle, n=self.n)

    def set_tool(self, tool) -> None:
        """vx_scene_set_tool: a vxsim.tool.Tool (or None to remove it); every
        later step advances it per substep after the prediction."""
        if tool is None:
            check(lib().vx_scene_set_tool(self.handle, None))
            return
        d, keep = tool.desc()
        check(lib().vx_scene_set_tool(self.handle, ctypes.byref(d)))
        del keep

    def tool_state(self) -> dict:
        from ._abi import ToolState
        st = ToolState()
        check(lib().vx_scene_tool_state(self.handle, ctypes.byref(st)))
        return dict(position=np.array(st.position[:]), rotation=np.array(st.rotation[:]),
                    velocity=np.array(st.velocity[:]), angular_velocity=np.array(st.angular_velocity[:]),
                    friction_heat=st.friction_heat, microsteps=st.microsteps, contacts=st.contacts,
                    candidates=st.candidates, status=st.status)

    def download_bond_forces(self) -> dict:
        return download_bond_forces(lib().vx_scene_download_bond_forces, self.handle, n=self.n)

    def export_rows(self) -> dict:
        na = self.active
        active = np.empty(na, np.uint32)
        nbr = np.empty((na, 6), np.uint32)
        blocks = np.empty((na, 54))
        diag = np.empty((na, 9))
        rhs = np.empty((na, 3))
        cnt = ctypes.c_uint32(0)
        check(lib().vx_scene_export_rows(self.handle, active.ctypes.data_as(c_u32p), nbr.ctypes.data_as(c_u32p),
                                         blocks.ctypes.data_as(c_dp), diag.ctypes.data_as(c_dp),
                                         rhs.ctypes.data_as(c_dp), ctypes.byref(cnt)))
        if cnt.value != na:
            raise SolverError(6, "active row count mismatch %d != %d" % (cnt.value, na))
        return dict(active=active, neighbors=nbr, blocks=blocks, diagonal=diag, rhs=rhs)

    def multigrid_level(self, level: int):
        """vx_scene_multigrid_level: level `level` (1 = the first coarse level)
        of the hierarchy the last multigrid solve built, as dict(dims,
        diagonal (n, 6), couplings (n, 3, 3, 3) [axis][row][column], masks
        (n,)); None past the coarsest level."""
        dims = np.zeros(3, np.uint32)
        check(lib().vx_scene_multigrid_level(self.handle, int(level), dims.ctypes.data_as(c_u32p), None, None, None))
        n = int(dims[0]) * int(dims[1]) * int(dims[2])
        if n == 0:
            return None
        d = np.empty((n, 6), np.float32)
        c = np.empty((n, 27), np.float32)
        m = np.empty(n, np.uint8)
        check(lib().vx_scene_multigrid_level(self.handle, int(level), dims.ctypes.data_as(c_u32p),
                                             d.ctypes.data_as(c_fp), c.ctypes.data_as(c_fp), m.ctypes.data_as(c_u8p)))
        return dict(dims=tuple(int(x) for x in dims), diagonal=d, couplings=c.reshape(n, 3, 3, 3), masks=m)

    def set_schedule(self, iterations_per_tick: int):
        """Inner PCG iterations per scheduling tick (graph replay)."""
        check(lib().vx_scene_set_schedule(self.handle, int(iterations_per_tick)))

    def set_shape(self, zcols: int = 8, deal: int = VX_DEAL_AUTO, update_threads: int = 512):
        """vx_scene_set_shape: the tile order of the inner kernels (z-column
        group, 0 or a power of two), their XCD deal (VX_DEAL_*) and the
        update kernel's threads per 512-row tile (512 or 256); results are
        identical under every shape."""
        check(lib().vx_scene_set_shape(self.handle, int(zcols), int(deal), int(update_threads)))

    def solve_timed(self, params=None, **kw) -> dict:
        """vx_scene_solve_timed: one solve with every bond SpMV and update
        dispatch timestamped; the report plus "spmv" / "update" statistics in
        microseconds per dispatch."""
        p = params if params is not None else solve_params(**kw)
        rep = SolveReport()
        times = KernelTimes()
        check(lib().vx_scene_solve_timed(self.handle, ctypes.byref(p), ctypes.byref(rep), ctypes.byref(times)))
        out = _report(rep)
        out["spmv"] = _report(times.spmv)
        out["update"] = _report(times.update)
        return out

    def time_spmv(self, reps=50) -> float:
        us = ctypes.c_double(0.0)
        check(lib().vx_scene_time_spmv(self.handle, int(reps), ctypes.byref(us)))
        return float(us.value)


def set_launch_mode(direct: bool) -> None:
    """vx_set_launch_mode: PCG chunks replayed as captured hipGraphs (the
    default) or launched kernel by kernel (direct), with identical results."""
    check(lib().vx_set_launch_mode(VX_LAUNCH_DIRECT if direct else VX_LAUNCH_GRAPHS))
