"""The oracle against the reference's mechanical fixtures and the laws its bond
forces obey.

* The axial fixture (src/experiments.rs:51-108) run by the oracle in the
  reference's own solver configuration (f32, block IC(0), 300 iterations,
  tolerance 1e-5) reproduces the recorded tension result
  (validation/mechanical-tests.txt:3: 1,022,767.250 N / 25,920.164 N, ratio
  39.458, first breaks 900/900) and the compression ratio of the latest
  checkpoint (validation/STATUS.md:14: 25.877).
* Bond forces and multipliers of a converged FP64 step satisfy the discrete
  equilibrium of every free degree of freedom (of REVs not bonded to a
  kinematic REV, whose bond has no last force),
  m (p - pred) = h^2 (load + sum over the REV's bonds of +-last_force),
  have |lambda| = h^2 |last_force| on undamaged bonds, and point as
  lambda = -F^T last_force h^2 in the (nearly unrotated) bond frame."""
import json
import os

import numpy as np
import pytest

import mech_fixture as M

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


@pytest.fixture(scope="module")
def fixtures_f32(O, scenes):
    return {label: M.axial_oracle(O, scenes, fib, sign, O.F32, O.IC0)
            for label, fib, sign in (("tension_par", 0, 1.0), ("tension_perp", 2, 1.0),
                                     ("compression_par", 0, -1.0), ("compression_perp", 2, -1.0))}


def test_tension_fixture_reproduces_the_reference_record(fixtures_f32):
    g = GOLDEN["tension_fixture"]
    (par, b1, _), (perp, b2, _) = fixtures_f32["tension_par"], fixtures_f32["tension_perp"]
    assert par == pytest.approx(g["parallel_N"], rel=1e-5)
    assert perp == pytest.approx(g["perpendicular_N"], rel=1e-5)
    assert [b1, b2] == g["first_breaks"]
    assert round(par / perp, 3) == g["ratio"]


def test_compression_fixture_reproduces_the_reference_ratio(fixtures_f32):
    (par, b1, _), (perp, b2, _) = fixtures_f32["compression_par"], fixtures_f32["compression_perp"]
    assert b1 > 0 and b2 > 0
    assert round(par / perp, 3) == GOLDEN["compression_fixture"]["ratio"]
    # the recorded acceptance criterion (src/experiments.rs:104-108)
    assert par / perp >= 10.0


def _loaded_step(O, scenes):
    s = scenes.large_64cube(dims=(8, 7, 6))
    cfg = s.config
    os_ = O.OracleScene(s)
    h = cfg.fixed_dt
    kin = s.kinematic != 0
    v = os_.velocities + np.array(cfg.gravity) * h
    pred = os_.positions.copy()
    pred[~kin] += v[~kin] * h
    rep = O.step(os_, cfg, O.F64, O.BJ, exact=True, exact_tolerance=1e-13, loads=s.loads, want_forces=True)
    return s, os_, pred, rep


def test_bond_forces_satisfy_equilibrium(O, scenes):
    s, os_, pred, rep = _loaded_step(O, scenes)
    h = s.config.fixed_dt
    lf = rep["last_force"]
    net = np.zeros((s.n, 3))
    for a in range(3):
        nb = s.neighbor[a]
        idx = np.nonzero(nb != 0xFFFFFFFF)[0]
        np.add.at(net, idx, lf[a, idx])
        np.subtract.at(net, nb[idx].astype(np.int64), lf[a, idx])
    free = s.free_mask()
    kin = s.kinematic != 0
    # A bond whose first REV is kinematic (the fixed face keeps its bonds)
    # has no last force: its second REV's sum is incomplete by definition.
    touched = np.zeros(s.n, bool)
    for a in range(3):
        nb = s.neighbor[a]
        idx = np.nonzero((nb != 0xFFFFFFFF) & kin)[0]
        touched[nb[idx].astype(np.int64)] = True
    m = s.mass.astype(np.float64)[:, None]
    lhs = m * (os_.positions - pred)
    rhs = h * h * (s.loads.astype(np.float64) + net)
    assert ((free != 0) & ~touched).sum() > 150
    scale = np.abs(h * h * lf).max()  # the largest single bond term
    for ax in range(3):
        sel = ((free >> ax & 1) == 1) & ~touched
        assert np.abs(lhs[sel, ax] - rhs[sel, ax]).max() <= 1e-6 * scale, ax
    # a kinematic REV's bonds carry no last force
    assert np.all(lf[:, kin] == 0.0)


def test_multipliers_are_scaled_rotated_forces(O, scenes):
    s, _, _, rep = _loaded_step(O, scenes)
    h = s.config.fixed_dt
    lf, lam = rep["last_force"], rep["lambda"]
    kin = s.kinematic != 0
    checked = 0
    for a in range(3):
        has = (s.neighbor[a] != 0xFFFFFFFF) & ~kin
        f, l = lf[a, has], lam[a, has]
        np.testing.assert_allclose(np.linalg.norm(l, axis=1), h * h * np.linalg.norm(f, axis=1), rtol=1e-10)
        # bond frame of an unrotated pair: columns e_a, e_a+1, e_a+2
        expect = -h * h * np.stack([f[:, a], f[:, (a + 1) % 3], f[:, (a + 2) % 3]], axis=1)
        assert np.abs(l - expect).max() <= 1e-3 * np.abs(expect).max()
        checked += int(has.sum())
    assert checked > 300
    # constraints with a kinematic first REV and a free second REV (the fixed
    # face keeps its bonds) still have multipliers
    nb = s.neighbor[2]
    edge = kin & (nb != 0xFFFFFFFF)
    edge[edge] = ~kin[nb[edge].astype(np.int64)]
    assert edge.any() and np.all(np.linalg.norm(lam[2, edge], axis=1) > 0)
