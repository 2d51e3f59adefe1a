"""Device contact projection (vx_contact_project) against the oracle's
collision::project (src/collision.rs:33-123): positions, previous positions,
contacts and counts bit for bit, on the reference's collision-test
configurations, on random lattice boxes with cracks, a fixed floor, fast
movers and targets, and on the 64^3 box; plus the boundary's error paths."""
import ctypes

import numpy as np
import pytest

from oracle import oracle
from vxsim import _abi, contact, lib, scenes
from vxsim.solver import SolverError

from test_oracle_contact import line_box

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[contact.AUTO, contact.WINDOWS], ids=["relaxation", "windows"], autouse=True)
def method(request):
    """Every test runs with the passes computed by relaxation (windows for a
    pass beyond its budget) and by windows alone."""
    contact.set_method(request.param)
    yield request.param
    contact.set_method(contact.AUTO)

CELL = np.float32(0.01)
EMPTY = 0xFFFFFFFF


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def check_parity(s, targets=None, label=""):
    print("contact parity", label, len(s["positions"]), flush=True)
    p, q, c, rep = contact.project(**s, targets=targets)
    po, qo, co, st = oracle.contact_project(**s, targets=targets)
    assert np.array_equal(bits(p), bits(po)), label
    assert np.array_equal(bits(q), bits(qo)), label
    for k in ("i", "j", "new_impact"):
        assert np.array_equal(c[k], co[k]), (label, k)
    assert np.array_equal(bits(c["normal"]), bits(co["normal"])), label
    assert rep["candidate_pairs"] == st["candidate_pairs"], label
    assert rep["pairs"] == st["pairs"], label
    assert rep["passes"] == st["passes"], label
    assert rep["contacts"] == st["contacts"], label
    return rep, st


def test_reference_configurations():
    s = line_box((3, 2, 1))
    s["bond_neighbor"][0, 0] = EMPTY
    s["positions"][1] = s["positions"][0] + np.float32(0.5) * np.array([CELL, 0, 0], np.float32)
    s["previous"] = s["positions"].copy()
    check_parity(s, label="crack")

    s = line_box((2, 1, 1))
    s["bond_neighbor"][0, 0] = EMPTY
    s["kinematic"] = np.array([False, True])
    s["positions"][0, 0] += np.float32(3.0) * CELL
    check_parity(s, label="tunnel")

    s = line_box((2, 1, 1))
    s["bond_neighbor"][:] = EMPTY
    s["positions"][1, 0] += np.float32(0.005)
    s["previous"][1] = s["positions"][1]
    s["positions"][0, 0] += np.float32(0.002503)
    s["positions"][1, 0] -= np.float32(0.002503)
    check_parity(s, targets=np.zeros((2, 3), bool), label="impact")

    s = line_box((3, 3, 3))
    c, sn = np.cos(0.37), np.sin(0.37)
    r = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1]])
    s["positions"] = (s["positions"].astype(np.float64) @ r.T).astype(np.float32)
    s["previous"] = s["positions"].copy()
    rep, _ = check_parity(s, label="rotated")
    assert rep["contacts"] == 0

    s = line_box((2, 1, 1))
    s["bond_neighbor"][:] = EMPTY
    s["positions"][:] = (s["positions"][0] + s["positions"][1]) * np.float32(0.5)
    s["previous"] = s["positions"].copy()
    check_parity(s, label="coincident")


@pytest.mark.parametrize("dims,seed,kw", [
    ((6, 5, 4), 1, {}),
    ((12, 10, 8), 3, {}),
    ((9, 7, 5), 4, {"fixed_floor": False, "jitter": 8e-4}),
    ((16, 12, 10), 5, {"movers": 0.05, "mover_reach": 3.0}),
    ((33, 7, 5), 6, {"broken": 0.3, "targeted": 0.2}),
    ((20, 20, 12), 8, {"jitter": 1.5e-3, "sag": 2e-4}),
])
def test_random_boxes_match_the_oracle(dims, seed, kw):
    s = scenes.contact_box(dims, seed=seed, **kw)
    targets = s.pop("targets")
    check_parity(s, targets=targets, label="%s seed %d" % (dims, seed))
    check_parity(s, targets=None, label="%s seed %d no targets" % (dims, seed))


def test_dense_pile_in_few_bins():
    # many movable voxels crowded into a few bins: long bin runs, long dependency chains
    rng = np.random.default_rng(11)
    n = 400
    s = {
        "positions": (rng.uniform(0, 3, (n, 3)) * CELL).astype(np.float32),
        "rest_positions": (np.stack([np.arange(n) * 3, np.zeros(n), np.zeros(n)], 1) * CELL).astype(np.float32),
        "mass": rng.uniform(1e-4, 1e-3, n).astype(np.float32),
        "kinematic": rng.uniform(size=n) < 0.1,
        "bond_neighbor": np.full((3, n), EMPTY, np.uint32),
    }
    s["previous"] = s["positions"].copy()
    rep, _ = check_parity(s, label="pile")
    assert rep["rounds"] + rep["relaxation_rounds"] > rep["passes"]


def test_64cube_matches_the_oracle(method):
    s = scenes.contact_box((64, 64, 64), seed=2)
    targets = s.pop("targets")
    rep, st = check_parity(s, targets=targets, label="64^3")
    assert st["pairs"] > 1_000_000 and st["contacts"] > 10_000
    if method == contact.WINDOWS:
        # the passes after the first took candidates only, and at least one was
        # undone and run again with more pairs: both paths are in this result
        assert rep["evaluated"] < 2 * st["pairs"] and rep["relaxed_passes"] == 0
        assert rep["redone"] >= 1
    else:
        # every pass reached its fixed point by relaxation, in a few rounds
        assert rep["relaxed_passes"] == rep["passes"] == st["passes"] and rep["rounds"] == 0
        assert rep["relaxation_rounds"] <= 16 * rep["passes"]


def raw_call(s, capacity):
    L = lib()
    n = len(s["positions"])
    pos = np.ascontiguousarray(s["positions"], np.float32).copy()
    prev = np.ascontiguousarray(s["previous"], np.float32).copy()
    rest = np.ascontiguousarray(s["rest_positions"], np.float32)
    mass = np.ascontiguousarray(s["mass"], np.float32)
    kin = np.ascontiguousarray(s["kinematic"], np.uint8)
    nb = np.ascontiguousarray(s["bond_neighbor"], np.uint32)
    out = (_abi.Contact * max(capacity, 1))()
    rep = _abi.ContactReport()
    st = L.vx_contact_project(n, pos.ctypes.data_as(_abi.c_fp), prev.ctypes.data_as(_abi.c_fp),
                              rest.ctypes.data_as(_abi.c_fp), mass.ctypes.data_as(_abi.c_fp),
                              kin.ctypes.data_as(_abi.c_u8p), None, nb.ctypes.data_as(_abi.c_u32p), out, capacity,
                              ctypes.byref(rep))
    return st, pos, prev, rep


def test_capacity_too_small_writes_nothing_and_reports_the_count():
    s = scenes.contact_box((10, 8, 6), seed=9)
    s.pop("targets")
    _, _, co, _ = oracle.contact_project(**s)
    assert len(co["i"]) > 1
    st, pos, prev, rep = raw_call(s, 1)
    assert st == _abi.VX_ERR_TOO_LARGE
    assert rep.contacts == len(co["i"])
    assert np.array_equal(bits(pos), bits(s["positions"]))
    assert np.array_equal(bits(prev), bits(s["previous"]))
    st, pos, prev, rep = raw_call(s, len(co["i"]))
    assert st == _abi.VX_OK


def same_or_both_nan(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    nan = np.isnan(a)
    return np.array_equal(nan, np.isnan(b)) and np.array_equal(bits(a[~nan]), bits(b[~nan]))


def test_nan_end_position_follows_the_reference():
    # f32::min / max skip a NaN operand, so the voxel is binned by its finite
    # start and the projection runs on, as in the reference
    s = line_box((3, 2, 2))
    s["bond_neighbor"][0, :] = EMPTY
    s["positions"][2, 1] = np.nan
    s["positions"][1, 0] -= np.float32(0.004)
    p, q, c, rep = contact.project(**s)
    po, qo, co, st = oracle.contact_project(**s)
    assert same_or_both_nan(p, po) and same_or_both_nan(q, qo)
    for k in ("i", "j", "new_impact"):
        assert np.array_equal(c[k], co[k])
    assert rep["contacts"] == st["contacts"] > 0


def test_invalid_inputs_are_rejected():
    s = line_box((3, 2, 2))
    s["positions"][2, 1] = np.nan
    s["previous"][2, 1] = np.nan  # no finite bound on that axis
    with pytest.raises(SolverError) as e:
        contact.project(**s)
    assert e.value.status == _abi.VX_ERR_INVALID_ARGUMENT
    s = line_box((3, 2, 2))
    s["positions"][0, 0] = np.float32(2.0e4)  # 2^21 cells away
    with pytest.raises(SolverError) as e:
        contact.project(**s)
    assert e.value.status == _abi.VX_ERR_INVALID_ARGUMENT
    s = line_box((3, 2, 2))
    s["bond_neighbor"][1, 0] = 12
    with pytest.raises(SolverError) as e:
        contact.project(**s)
    assert e.value.status == _abi.VX_ERR_INVALID_ARGUMENT


def test_empty_and_all_kinematic_scenes():
    st, _, _, rep = raw_call({"positions": np.zeros((0, 3)), "previous": np.zeros((0, 3)),
                              "rest_positions": np.zeros((0, 3)), "mass": np.zeros(0),
                              "kinematic": np.zeros(0, bool), "bond_neighbor": np.zeros((3, 0))}, 0)
    assert st == _abi.VX_OK and rep.contacts == 0 and rep.passes == 1
    s = line_box((4, 3, 2), kinematic=np.ones(24, bool))
    s["bond_neighbor"][:] = EMPTY
    rep, _ = check_parity(s, label="all kinematic")
    assert rep["pairs"] == 0 and rep["contacts"] == 0


def test_repeated_calls_reuse_buffers_and_agree():
    big = scenes.contact_box((24, 20, 16), seed=12)
    small = scenes.contact_box((7, 6, 5), seed=13)
    for s in (big, small, big):
        t = s["targets"]
        s2 = {k: v for k, v in s.items() if k != "targets"}
        check_parity(s2, targets=t, label="sequence %d" % len(s2["positions"]))


def test_state_left_by_the_blade_cut_matches_the_oracle():
    # the projection as it follows the solve: the state a device step with
    # fracture leaves (deformed lattice, broken bonds, the blade's targets)
    from vxsim import grid
    blade = (10, 5, 8)
    sc = scenes.wood_block_blade(dims=(32, 32, 20), blade=blade)
    cfg = sc.config
    sp = grid.solve_params(tolerance=1e-6, inner_tolerance=3e-3, max_inner=20000, max_outer=12)
    with grid.GridScene(sc) as gs:
        for k in range(3):
            start = gs.download()["positions"]
            tg = scenes.blade_targets(sc, blade, 5.0, cfg.fixed_dt, k)
            gs.set_targets_and_loads(tg, sc.loads)
            gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                    corotation_iterations=cfg.corotation_iterations, corotation_tolerance=cfg.corotation_tolerance,
                    solve=sp, fracture=scenes.FractureConfig())
        end = gs.download()["positions"]
        bonds = gs.download_bonds()["neighbor"]
    s = {"positions": end.astype(np.float32), "previous": start.astype(np.float32),
         "rest_positions": sc.positions.astype(np.float32), "mass": sc.mass.astype(np.float32),
         "kinematic": sc.kinematic != 0, "bond_neighbor": bonds}
    assert np.count_nonzero(bonds == EMPTY) > np.count_nonzero(sc.neighbor == EMPTY)  # bonds broke
    rep, st = check_parity(s, targets=~np.isnan(tg), label="blade cut state")
    assert st["contacts"] > 0


def test_negative_zero_coordinates_keep_the_reference_bits():
    # separate adds exact zeros; only a -0 coordinate can tell +0 from -0
    # apart (it turns into +0), so plant -0 in positions and previous ones
    s = scenes.contact_box((10, 8, 6), seed=21, jitter=8e-4)
    targets = s.pop("targets")
    rng = np.random.default_rng(21)
    movable = np.nonzero(~s["kinematic"])[0]
    for key in ("positions", "previous"):
        pick = rng.choice(movable, size=60, replace=False)
        comp = rng.integers(0, 3, size=60)
        s[key][pick, comp] = np.float32(-0.0)
    assert np.count_nonzero(np.signbit(s["positions"]) & (s["positions"] == 0)) > 0
    rep, st = check_parity(s, targets=targets, label="-0 coordinates")
    p, _, _, _ = contact.project(**s, targets=targets)
    po, _, _, _ = oracle.contact_project(**s, targets=targets)
    zero = (po == 0)
    assert np.array_equal(np.signbit(p[zero]), np.signbit(po[zero]))


def test_negative_zero_on_kinematic_voxels_next_to_movers():
    """Kinematic endpoints receive separate's zero-weight share too
    (src/collision.rs:657-663): the higher-index endpoint j gets normal * +0,
    which turns a -0 coordinate into +0, the lower one keeps -0. With the
    fixed floor (lowest indices) and the top layer kinematic, 20% of the
    bonds broken, plant -0 in X or Y of the kinematic voxels nearest the
    box's centre lines (half a cell from their place) and compare every bit,
    signs of zero included, with the oracle; some of them flip."""
    s = scenes.contact_box((10, 8, 6), seed=23, jitter=8e-4, movers=0.05, broken=0.2)
    targets = s.pop("targets")
    rest = s["rest_positions"]
    top = rest[:, 2] == rest[:, 2].max()
    s["kinematic"] = s["kinematic"] | top
    targets[top] = False
    kin = np.nonzero(s["kinematic"])[0]
    planted = []
    for axis in (0, 1):
        near = kin[np.abs(rest[kin, axis]) == np.float32(0.005)]
        for key in ("positions", "previous"):
            s[key][near, axis] = np.float32(-0.0)
        planted.append((near, axis))
    rep, st = check_parity(s, targets=targets, label="-0 on kinematic voxels")
    p, q, _, _ = contact.project(**s, targets=targets)
    po, qo, _, _ = oracle.contact_project(**s, targets=targets)
    flipped = 0
    for near, axis in planted:
        for a, b in ((p, po), (q, qo)):
            assert np.array_equal(np.signbit(a[near, axis]), np.signbit(b[near, axis]))
        flipped += np.count_nonzero(~np.signbit(po[near, axis]))
    assert flipped > 0


def test_quiet_state_is_settled_in_one_parallel_pass():
    """A state the projection leaves unmoved (a cracked lattice box projected
    twice by the oracle: one pass, touching contacts only) is settled by
    evaluating every pair once, all at once: no window runs, and positions,
    previous positions and contacts equal the oracle's bit for bit."""
    s = scenes.contact_box((24, 24, 16), seed=4, jitter=0.0, sag=0.0, movers=0.0, broken=0.1)
    targets = s.pop("targets")
    for _ in range(2):
        po, qo, _, _ = oracle.contact_project(**s, targets=targets)
        s["positions"], s["previous"] = po, qo
    rep, st = check_parity(s, targets=targets, label="quiet state")
    assert st["passes"] == 1 and st["contacts"] > 1000
    assert rep["rounds"] == 0 and rep["evaluated"] == rep["pairs"]
