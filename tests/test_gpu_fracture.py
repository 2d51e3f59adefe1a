"""GPU parity of the fracture pass (src/solver.rs:627-665 with
fracture::update, src/fracture.rs:210-292) against the FP64 oracle, through
the C ABI: the set of broken bonds is identical, the dissipated energies and
the bond history (damage, plastic strain, equivalent plastic strain, plastic
work, fatigue) agree to the solve's accuracy, and the steps after a break
(assembled without the broken bonds, with the plastic rest offsets) keep
matching. Bond state is stored in float32, the format of the reference's
Bond, on the device and (rounded between steps) in the oracle."""
import numpy as np
import pytest

from vxsim.scenes import FractureConfig, HISTORY_FIELDS

pytestmark = pytest.mark.gpu
EMPTY = 0xFFFFFFFF
STEPS = 3
ENERGIES = ("plastic_heat", "stored_plastic", "surface_work", "damage_heat", "released")


def run_device(s, steps, fracture=FractureConfig(), tol=1e-11, preconditioner=0):
    from vxsim import grid
    cfg = s.config
    sp = grid.solve_params(tolerance=tol, inner_tolerance=1e-3, max_inner=20000, preconditioner=preconditioner)
    reps = []
    with grid.GridScene(s) as gs:
        for _ in range(steps):
            reps.append(gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                                corotation_iterations=cfg.corotation_iterations,
                                corotation_tolerance=cfg.corotation_tolerance, solve=sp, fracture=fracture))
        return reps, gs.download(), gs.download_bonds()


def run_oracle(O, s, steps):
    """FP64 oracle steps; between steps the bond state is rounded to float32,
    the storage format of the reference's Bond and of the device planes."""
    os_ = O.OracleScene(s)
    hist = O.FractureHistory(s.n)
    reps = []
    for _ in range(steps):
        reps.append(O.step_fracture(os_, s.config, FractureConfig(), hist, s.strength, O.F64, exact=True,
                                    exact_tolerance=1e-12, targets=s.targets, loads=s.loads))
        for a in [os_.damage, os_.plastic_strain] + [getattr(hist, k) for k in HISTORY_FIELDS]:
            a[...] = a.astype(np.float32)
    return reps, os_, hist


def check_parity(O, s, steps=STEPS, preconditioner=0):
    reps_d, out, bonds = run_device(s, steps, preconditioner=preconditioner)
    reps_o, os_, hist = run_oracle(O, s, steps)
    broken_total = 0
    # Energies against the run's energy scale: a later step's small residual
    # flow is a difference of nearly equal bond energies.
    scale = max(max(sum(ro["dissipated"].values()) for ro in reps_o), 1e-30)
    for k, (rd, ro) in enumerate(zip(reps_d, reps_o)):
        fr = rd["fracture"]
        assert fr["seconds"] > 0.0  # the pass ran (HIP-event timed)
        assert fr["bonds_broken"] == ro["broken"], (s.name, k)
        broken_total += ro["broken"]
        for e in ENERGIES:
            assert abs(fr[e] - ro["dissipated"][e]) <= 1e-6 * scale + 1e-12, (s.name, k, e, fr[e], ro["dissipated"][e])
    # the same bonds are gone
    assert np.array_equal(bonds["neighbor"], os_.neighbor), s.name
    # bond history to the solve's accuracy (float32 storage on the device)
    intact = os_.neighbor != EMPTY
    np.testing.assert_allclose(bonds["damage"], os_.damage, rtol=1e-5, atol=1e-7)
    pl_o = os_.plastic_strain
    np.testing.assert_allclose(bonds["plastic_strain"], pl_o, rtol=1e-5, atol=1e-6 * max(np.abs(pl_o).max(), 1e-12))
    ratio_scale = np.abs(hist.previous_stress_ratio).max()
    for f, name in enumerate(HISTORY_FIELDS):
        ref = getattr(hist, name)
        got = bonds["history"][..., f]
        # The stress-ratio fields (and the increment, a difference of two
        # ratios) carry the ratio's absolute error.
        scale = ratio_scale if name.endswith(("stress_ratio", "increment")) else np.abs(ref).max()
        np.testing.assert_allclose(got[intact], ref[intact], rtol=1e-5, atol=1e-6 * max(scale, 1e-12), err_msg=name)
    # positions after the breaks (the later steps run without the broken bonds)
    start = s.positions.astype(np.float64)
    disp_o = os_.positions - start
    disp = out["positions"] - start
    assert np.linalg.norm(disp - disp_o) <= 1e-6 * np.linalg.norm(disp_o), s.name
    return broken_total, hist


@pytest.mark.parametrize("material,pull", [("stone", 2e-6), ("stone", 5e-5), ("iron", 1e-3), ("iron", 8e-3)])
def test_pulled_column_matches_oracle(gpu, O, scenes, material, pull):
    s = scenes.pulled_column(material, pull)
    broken, hist = check_parity(O, s)
    if material == "stone":
        assert broken == (45 if pull > 1e-5 else 0)
    else:
        assert broken == 0 and hist.equivalent_plastic_strain.max() > 0.0


def test_fracture_every_substep_matches_oracle(gpu, O, scenes):
    """Four substeps per step (the reference's default SceneConfig): the pass
    runs after each substep's corotation passes, with the substep length as
    its time step; later substeps assemble without the bonds broken before."""
    s = scenes.pulled_mixed((10, 10, 8), 9, 3e-4)
    s.config.substeps = 4
    broken, hist = check_parity(O, s, steps=2)
    assert broken > 50


# vx_solve_params.preconditioner: 0 block-Jacobi, 1 red-black reduced system,
# 2 aggregation multigrid (its hierarchy rebuilt after every assembly, so
# after the bonds a step breaks)
PRECONDITIONERS = {"block_jacobi": 0, "red_black": 1, "multigrid": 2}


@pytest.mark.parametrize("precond", list(PRECONDITIONERS))
def test_pulled_mixed_box_matches_oracle(gpu, O, scenes, precond):
    """Brittle breaks, cross-grain wood breaks and iron / along-grain wood
    plastic flow in one 12^3 box, three steps; the same broken-bond set with
    every preconditioner."""
    s = scenes.pulled_mixed((12, 12, 12), 4, 4e-4)
    broken, hist = check_parity(O, s, preconditioner=PRECONDITIONERS[precond])
    assert broken > 500 and np.count_nonzero(hist.equivalent_plastic_strain) > 100


def exempt(os_, hist, bonds):
    """Bonds broken on one side only, each within 1e-5 of a failure
    threshold in the oracle's history (stress ratio, damage or fatigue at
    one): the reference's 1e-5 exemption for broken-bond sets."""
    diff = np.argwhere((bonds["neighbor"] == EMPTY) != (os_.neighbor == EMPTY))
    for a, v in diff:
        near = (abs(abs(hist.previous_stress_ratio[a, v]) - 1.0) <= 1e-5 or abs(os_.damage[a, v] - 1.0) <= 1e-5
                or abs(hist.fatigue_damage[a, v] - 1.0) <= 1e-5)
        assert near, ("bond broken on one side only", int(a), int(v), hist.previous_stress_ratio[a, v],
                      os_.damage[a, v])
    return len(diff)


def blade_cut_parity(O, scenes, dims, blade, steps, precond, min_broken):
    """The cutting block with per-step axis targets through
    vx_scene_set_targets_and_loads, six corotation passes and fracture every
    step, against the FP64 oracle: the broken-bond set identical up to the
    1e-5 exemption; energies, damage and positions to the solve's accuracy."""
    from vxsim import grid
    s = scenes.wood_block_blade(dims=dims, blade=blade)
    cfg = s.config
    sp = grid.solve_params(tolerance=1e-11, inner_tolerance=1e-3, max_inner=40000,
                           preconditioner=PRECONDITIONERS[precond])
    reps_d = []
    with grid.GridScene(s) as gs:
        for k in range(steps):
            gs.set_targets_and_loads(scenes.blade_targets(s, blade, 5.0, cfg.fixed_dt, k), s.loads)
            reps_d.append(gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                                  corotation_iterations=cfg.corotation_iterations,
                                  corotation_tolerance=cfg.corotation_tolerance, solve=sp,
                                  fracture=FractureConfig()))
        out, bonds = gs.download(), gs.download_bonds()
    os_ = O.OracleScene(s)
    hist = O.FractureHistory(s.n)
    reps_o = []
    for k in range(steps):
        reps_o.append(O.step_fracture(os_, cfg, FractureConfig(), hist, s.strength, O.F64, exact=True,
                                      exact_tolerance=1e-12, targets=scenes.blade_targets(s, blade, 5.0, cfg.fixed_dt, k),
                                      loads=s.loads))
        for arr in [os_.damage, os_.plastic_strain] + [getattr(hist, f) for f in HISTORY_FIELDS]:
            arr[...] = arr.astype(np.float32)
    assert sum(r["broken"] for r in reps_o) > min_broken
    differing = exempt(os_, hist, bonds)
    scale = max(sum(r["dissipated"].values()) for r in reps_o)
    for rd, ro in zip(reps_d, reps_o):
        assert abs(int(rd["fracture"]["bonds_broken"]) - ro["broken"]) <= differing
        if differing == 0:
            for e in ENERGIES:
                assert abs(rd["fracture"][e] - ro["dissipated"][e]) <= 1e-6 * scale, e
    both = (bonds["neighbor"] != EMPTY) & (os_.neighbor != EMPTY)
    np.testing.assert_allclose(bonds["damage"][both], os_.damage[both], rtol=1e-5, atol=1e-6)
    start = s.positions.astype(np.float64)
    disp_o = os_.positions - start
    assert np.linalg.norm(out["positions"] - start - disp_o) <= 1e-6 * np.linalg.norm(disp_o)
    return [r["broken"] for r in reps_o], differing


@pytest.mark.parametrize("precond", list(PRECONDITIONERS))
def test_blade_cut_matches_oracle(gpu, O, scenes, precond):
    """The cutting block at reduced size: 24x24x16 wood, grain Z, blade edge
    8x3 REVs driven 5 mm per 1 ms step, two steps."""
    blade_cut_parity(O, scenes, (24, 24, 16), (8, 3, 5), 2, precond, 1000)


def test_blade_cut_64x64x32_matches_oracle(gpu, O, scenes):
    """The cut at a larger size with the full 40x5 blade edge: 64x64x32 wood
    (131,072 REVs), three steps (about 36,000 bonds broken)."""
    broken, differing = blade_cut_parity(O, scenes, (64, 64, 32), (40, 5, 20), 3, "block_jacobi", 30000)
    print("64x64x32 cut: bonds broken per step %s, exempt differences %d" % (broken, differing))


def test_targets_and_loads_update(gpu, scenes):
    """Per-step targets reach the solve: a step after moving the targets
    equals a scene created with those targets (warm start off), and NaN-only
    targets free every non-kinematic axis."""
    from vxsim import grid
    from vxsim.solver import SolverError
    s = scenes.pulled_column("stone", 2e-6)
    moved = s.copy()
    moved.targets = s.targets.copy()
    moved.targets[~np.isnan(moved.targets[:, 2]), 2] += np.float32(1e-6)
    moved.loads = s.loads.copy()
    moved.loads[:, 0] = np.float32(0.5)
    kw = dict(fixed_dt=0.001, gravity=(0.0, 0.0, 0.0), corotation_iterations=2,
              solve=grid.solve_params(tolerance=1e-10))
    with grid.GridScene(s) as a, grid.GridScene(moved) as b:
        a.set_targets_and_loads(moved.targets, moved.loads)
        assert a.active == b.active
        a.step(**kw)
        b.step(**kw)
        assert np.array_equal(a.download()["positions"], b.download()["positions"])
        a.set_targets_and_loads(None, None)
        assert a.active == s.n
        bad = moved.targets.copy()
        bad[0, 0] = np.inf
        with pytest.raises(SolverError):
            a.set_targets_and_loads(bad, None)


def test_fracture_off_leaves_bonds_alone(gpu, scenes):
    """Without the pass the step is the plain step: same positions bit for
    bit, no bond touched, nothing reported."""
    s = scenes.pulled_column("stone", 5e-5)
    reps, out, bonds = run_device(s, 2, fracture=None)
    plain = s.copy()
    plain.strength = None
    reps_p, out_p, _ = run_device(plain, 2, fracture=None)
    assert np.array_equal(out["positions"], out_p["positions"])
    assert np.array_equal(bonds["neighbor"], s.neighbor)
    assert not bonds["damage"].any() and not bonds["history"].any()
    assert all(r["fracture"]["bonds_broken"] == 0 and r["fracture"]["surface_work"] == 0.0 and
               r["fracture"]["seconds"] == 0.0 for r in reps)


def test_fracture_needs_strength(gpu, scenes):
    from vxsim import grid
    from vxsim.solver import SolverError
    s = scenes.pulled_column("stone", 5e-5)
    s.strength = None
    with grid.GridScene(s) as gs:
        with pytest.raises(SolverError):
            gs.step(fixed_dt=0.001, gravity=(0, 0, 0), fracture=FractureConfig())


def test_batched_fracture_same_alone_and_in_mix(gpu, scenes):
    """Scenes with and without the fracture pass share one arena; each one's
    positions and bonds are bit-identical to the scene stepped alone."""
    from vxsim import batch, grid
    cases = [(scenes.pulled_mixed((10, 9, 11), 5, 4e-4), FractureConfig()),
             (scenes.pulled_column("stone", 5e-5), FractureConfig()),
             (scenes.pulled_mixed((8, 12, 7), 6, 4e-4), None),
             (scenes.pulled_column("iron", 8e-3, dims=(4, 3, 9)), FractureConfig())]
    sp = grid.solve_params(tolerance=1e-9, inner_tolerance=1e-3, max_inner=20000)
    params = []
    with batch.Batch(capacity_rows=16 * 512, max_scenes=8) as B:
        slots = []
        for s, f in cases:
            c = s.config
            p = grid.step_params(c.fixed_dt, c.gravity, c.substeps, c.corotation_iterations,
                                 c.corotation_tolerance, sp, fracture=f)
            params.append(p)
            slots.append(B.admit(s, 2, p))
        B.run()
        got = [(B.download(k), B.download_bonds(k), B.report(k)) for k in slots]
    for (s, f), (out, bonds, rep) in zip(cases, got):
        assert rep["state"] == batch.DONE
        c = s.config
        with grid.GridScene(s) as gs:
            total = 0
            for _ in range(2):
                r = gs.step(fixed_dt=c.fixed_dt, gravity=c.gravity, substeps=c.substeps,
                            corotation_iterations=c.corotation_iterations,
                            corotation_tolerance=c.corotation_tolerance, solve=sp, fracture=f)
                total += r["fracture"]["bonds_broken"]
            ref, ref_b = gs.download(), gs.download_bonds()
        assert np.array_equal(out["positions"], ref["positions"]), s.name
        for key in ("neighbor", "damage", "plastic_strain", "history"):
            assert np.array_equal(bonds[key], ref_b[key]), (s.name, key)
        assert rep["fracture"]["bonds_broken"] == total
        if f is None:
            assert total == 0
