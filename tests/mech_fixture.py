"""The reference's axial loading fixture (src/experiments.rs:14-108), shared by
the oracle test and the device test: a 10x10x10 specimen of the built-in wood
(scenes/wood_chop.ron with no branch or ground, fibres along X or Z, no
gravity, one 1 ms substep, 300 solver iterations, tolerance 1e-5), moved 1 m up
in Z; the REVs of its -X face held at rest, those of its +X face displaced
along X in a slow monotonic ramp (0.12 of the specimen length over 500 steps),
stopped at the first step that breaks a bond. The reaction is |x| of the sum
of the last forces of the bonds touching the +X grip (src/experiments.rs:
32-49). Returns (peak reaction, bonds broken in the first breaking step)."""
import numpy as np

CELL = np.float32(0.01)


def specimen(scenes, fiber_axis: int):
    s = scenes.uniform_box((10, 10, 10), "wood_builtin", grain=fiber_axis, anisotropy=1.0)
    s.positions[:, 2] += np.float32(1.0)
    s.config = scenes.StepConfig(fixed_dt=0.001, substeps=1, iterations=300, corotation_iterations=4,
                                 solver_relative_tolerance=1e-5, corotation_tolerance=1e-5, gravity=(0.0, 0.0, 0.0))
    return s


def grips(s):
    rest = s.positions
    lo, hi = rest[:, 0].min(), rest[:, 0].max()
    return lo, hi, rest[:, 0] < lo + CELL * np.float32(0.1), rest[:, 0] > hi - CELL * np.float32(0.1)


def ramp_targets(s, sign: float, n: int):
    """Targets of ramp step n (1-based), in the reference's f32 order."""
    lo, hi, left, right = grips(s)
    disp = np.float32(np.float32(np.float32(np.float32(sign) * (hi - lo)) * np.float32(0.12)) * np.float32(n))
    disp = np.float32(disp / np.float32(500.0))
    t = np.full((s.n, 3), np.nan, np.float32)
    t[left] = s.positions[left]
    t[right] = s.positions[right]
    t[right, 0] = s.positions[right, 0] + disp
    return t


def reaction(last_force, rest_neighbor, right) -> float:
    """|x| of the sum over bonds of the last force, + where the bond's first
    REV is in the grip, - where its second is (src/experiments.rs:32-49)."""
    f = 0.0
    for a in range(3):
        nb = rest_neighbor[a]
        has = nb != 0xFFFFFFFF
        idx = np.nonzero(has)[0]
        fx = last_force[a, idx, 0]
        f += fx[right[idx]].sum() - fx[right[nb[idx]]].sum()
    return abs(f)


def axial_oracle(O, scenes, fiber_axis: int, sign: float, mode, precond, exact=False, max_steps=500):
    s = specimen(scenes, fiber_axis)
    rest_neighbor = s.neighbor.copy()
    _, _, _, right = grips(s)
    os_ = O.OracleScene(s)
    hist = O.FractureHistory(s.n)
    peak = 0.0
    for n in range(1, max_steps + 1):
        rep = O.step_fracture(os_, s.config, scenes.FractureConfig(), hist, s.strength, mode, precond, exact=exact,
                              exact_tolerance=1e-12, targets=ramp_targets(s, sign, n), want_forces=True)
        peak = max(peak, reaction(rep["last_force"], rest_neighbor, right))
        if rep["broken"] > 0:
            return peak, int(rep["broken"]), n
    return peak, 0, max_steps


def axial_device(scenes, fiber_axis: int, sign: float, tolerance=1e-10, max_steps=500):
    from vxsim import grid
    s = specimen(scenes, fiber_axis)
    rest_neighbor = s.neighbor.copy()
    _, _, _, right = grips(s)
    cfg = s.config
    peak = 0.0
    broken_total = 0
    with grid.GridScene(s) as gs:
        for n in range(1, max_steps + 1):
            gs.set_targets_and_loads(ramp_targets(s, sign, n), s.loads)
            rep = gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=1,
                          corotation_iterations=cfg.corotation_iterations, corotation_tolerance=cfg.corotation_tolerance,
                          solve=grid.solve_params(tolerance=tolerance, max_inner=20000),
                          fracture=scenes.FractureConfig())
            lf = gs.download_bond_forces()["last_force"]
            peak = max(peak, reaction(lf, rest_neighbor, right))
            broken_total = int(rep["fracture"]["bonds_broken"])
            if broken_total > 0:
                return peak, broken_total, n
    return peak, 0, max_steps
