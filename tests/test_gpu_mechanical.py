"""GPU parity of the bond forces and multipliers (the reference's
BondWorkspace::last_force, src/solver.rs:627-640, and ::lambda,
src/solver.rs:1824-1827) against the FP64 oracle, with and without the
fracture pass, and the reference's axial loading fixture
(src/experiments.rs:51-108) run on the device: the peak reactions and the
first breaking step against the oracle's FP64 run, and the recorded
results of the reference (validation/mechanical-tests.txt:3,
validation/STATUS.md:13-14)."""
import json
import os

import numpy as np
import pytest

import mech_fixture as M
from vxsim.scenes import FractureConfig, HISTORY_FIELDS

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


def _device_forces(s, steps, fracture):
    from vxsim import grid
    cfg = s.config
    sp = grid.solve_params(tolerance=1e-11, inner_tolerance=1e-3, max_inner=20000)
    with grid.GridScene(s) as gs:
        for _ in range(steps):
            rep = gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                          corotation_iterations=cfg.corotation_iterations,
                          corotation_tolerance=cfg.corotation_tolerance, solve=sp, fracture=fracture)
        return rep, gs.download_bond_forces()


def _close(got, ref, what):
    scale = np.abs(ref).max()
    assert scale > 0, what
    assert np.abs(got - ref).max() <= 1e-6 * scale, (what, np.abs(got - ref).max(), scale)
    # zero exactly where there is no constraint (or a kinematic first REV)
    assert np.array_equal(got == 0.0, ref == 0.0), what


def test_bond_forces_match_oracle(gpu, O, scenes):
    s = scenes.large_64cube(dims=(12, 11, 10))
    _, f = _device_forces(s, 1, None)
    os_ = O.OracleScene(s)
    rep = O.step(os_, s.config, O.F64, O.BJ, exact=True, exact_tolerance=1e-13, loads=s.loads, want_forces=True)
    _close(f["last_force"], rep["last_force"], "last_force")
    _close(f["lambda"], rep["lambda"], "lambda")


def test_bond_forces_before_fracture_update_match_oracle(gpu, O, scenes):
    """A pulled mixed box whose second step breaks bonds: the forces of that
    step are taken before the fracture pass updates (and removes) them."""
    s = scenes.pulled_mixed((10, 9, 8), seed=5, pull=2e-4)  # step 1 breaks none, step 2 eight bonds
    rep_d, f = _device_forces(s, 2, FractureConfig())
    # the oracle's two steps (bond state rounded to float32 between them),
    # the second with the forces
    os_ = O.OracleScene(s)
    hist = O.FractureHistory(s.n)
    for k in range(2):
        r = O.step_fracture(os_, s.config, FractureConfig(), hist, s.strength, O.F64, exact=True,
                            exact_tolerance=1e-12, targets=s.targets, loads=s.loads, want_forces=(k == 1))
        for a in [os_.damage, os_.plastic_strain] + [getattr(hist, n) for n in HISTORY_FIELDS]:
            a[...] = a.astype(np.float32)
    assert r["broken"] == rep_d["fracture"]["bonds_broken"] > 0
    _close(f["last_force"], r["last_force"], "last_force")
    _close(f["lambda"], r["lambda"], "lambda")


@pytest.fixture(scope="module")
def fixtures(O, scenes):
    out = {}
    for label, fib, sign in (("tension_par", 0, 1.0), ("tension_perp", 2, 1.0), ("compression_par", 0, -1.0),
                             ("compression_perp", 2, -1.0)):
        out[label] = (M.axial_device(scenes, fib, sign), M.axial_oracle(O, scenes, fib, sign, O.F64, O.BJ, exact=True))
    return out


def test_axial_fixtures_match_fp64_oracle(gpu, fixtures):
    for label, ((pd, bd, nd), (po, bo, no)) in fixtures.items():
        assert (bd, nd) == (bo, no), label  # the same first breaking step and bond count
        assert pd == pytest.approx(po, rel=1e-6), label


def test_axial_fixtures_match_reference_records(gpu, fixtures):
    """The reference ran the fixture with its f32 CPU solve to a relative
    residual of 1e-5; the FP64 solution differs from it by about 1.3e-5 in
    the peak reaction (the oracle in the reference's own configuration
    reproduces the record to 1e-6), so the device's FP64-refined peaks are
    held to 5e-5 of the record. Break counts and the ratios' acceptance
    (>= 10) are exact."""
    g = GOLDEN["tension_fixture"]
    (par, b1, _), _ = fixtures["tension_par"]
    (perp, b2, _), _ = fixtures["tension_perp"]
    assert par == pytest.approx(g["parallel_N"], rel=5e-5)
    assert perp == pytest.approx(g["perpendicular_N"], rel=5e-5)
    assert [b1, b2] == g["first_breaks"]
    assert par / perp == pytest.approx(g["ratio"], abs=2e-3) and par / perp >= 10.0
    (cpar, c1, _), _ = fixtures["compression_par"]
    (cperp, c2, _), _ = fixtures["compression_perp"]
    assert c1 > 0 and c2 > 0
    assert cpar / cperp == pytest.approx(GOLDEN["compression_fixture"]["ratio"], abs=2e-3) and cpar / cperp >= 10.0
