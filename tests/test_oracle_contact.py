"""The contact-projection oracle (collision::project, src/collision.rs:33-123)
pinned by the reference's own collision tests (src/collision.rs:667-1016),
restated on the same configurations, by a brute-force broad phase and by
the momentum balance of `separate`."""
import numpy as np
import pytest

from oracle import oracle
from vxsim import scenes

CELL = np.float32(0.01)
EMPTY = 0xFFFFFFFF


def line_box(dims, kinematic=None):
    """A lattice box as the reference's Scene::new lays it out (centres at
    key*CELL + CELL/2, keys centred in x and y), with its lattice bonds."""
    nx, ny, nz = dims
    n = nx * ny * nz
    idx = np.arange(n)
    key = np.stack([idx % nx - nx // 2, (idx // nx) % ny - ny // 2, idx // (nx * ny)], axis=1)
    rest = (key.astype(np.float32) * CELL + CELL / np.float32(2)).astype(np.float32)
    kin = np.zeros(n, bool) if kinematic is None else np.asarray(kinematic, bool)
    nb = scenes.lattice_topology(dims, kin)
    return {"positions": rest.copy(), "previous": rest.copy(), "rest_positions": rest,
            "mass": np.full(n, 6e-4, np.float32), "kinematic": kin, "bond_neighbor": nb}


def test_local_exclusion_does_not_hide_a_crack_or_remote_self_contact():
    # src/collision.rs:690-710, a 3 x 2 x 1 box
    s = line_box((3, 2, 1))
    assert oracle.local_continuum_pair(s["rest_positions"], s["bond_neighbor"], 0, 4)
    assert not oracle.local_continuum_pair(s["rest_positions"], s["bond_neighbor"], 0, 2)
    s["bond_neighbor"][0, 0] = EMPTY
    assert not oracle.local_continuum_pair(s["rest_positions"], s["bond_neighbor"], 0, 1)
    s["positions"][1] = s["positions"][0] + np.float32(0.5) * np.array([CELL, 0, 0], np.float32)
    s["previous"] = s["positions"].copy()
    _, _, c, _ = oracle.contact_project(**s)
    assert any(int(i) == 0 and int(j) == 1 for i, j in zip(c["i"], c["j"]))


def test_moving_voxel_cannot_tunnel_through_fixed_voxel():
    # src/collision.rs:731-747: voxel 1 kinematic, voxel 0 jumps 3 cells through it
    s = line_box((2, 1, 1))
    s["bond_neighbor"][0, 0] = EMPTY
    s["kinematic"] = np.array([False, True])
    wall = s["positions"][1].copy()
    s["positions"][0, 0] += np.float32(3.0) * CELL
    p, _, c, _ = oracle.contact_project(**s)
    assert len(c["i"]) > 0
    assert np.array_equal(p[1], wall)
    assert wall[0] - p[0, 0] >= CELL - np.float32(1e-6)


def test_touching_contact_persists_after_projection_slop():
    # src/collision.rs:910-938: an impact, then the resting contact it leaves
    s = line_box((2, 1, 1))
    s["bond_neighbor"][:] = EMPTY  # two separate spawned voxels
    s["mass"][:] = np.float32(7870.0) * CELL ** 3
    s["positions"][1, 0] += np.float32(0.005)
    s["previous"][1] = s["positions"][1]
    s["positions"][0, 0] += np.float32(0.002503)
    s["positions"][1, 0] -= np.float32(0.002503)
    targets = np.zeros((2, 3), bool)
    p, q, c, _ = oracle.contact_project(**s, targets=targets)
    assert len(c["i"]) == 1 and bool(c["new_impact"][0])
    s["positions"], s["previous"] = p, p.copy()
    p2, _, c2, _ = oracle.contact_project(**s, targets=targets)
    assert len(c2["i"]) == 1 and not bool(c2["new_impact"][0])
    assert np.linalg.norm(p2[1].astype(np.float64) - p2[0]) <= 0.01 + 2e-6


def test_rigidly_rotated_continuum_has_no_internal_box_contacts():
    # src/collision.rs:673-688: a 3^3 box rotated 0.37 rad about z
    s = line_box((3, 3, 3))
    c, sn = np.cos(0.37), np.sin(0.37)
    r = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1]])
    s["positions"] = (s["positions"].astype(np.float64) @ r.T).astype(np.float32)
    s["previous"] = s["positions"].copy()
    before = s["positions"].copy()
    p, _, con, st = oracle.contact_project(**s)
    assert len(con["i"]) == 0
    assert np.array_equal(p.view(np.uint32), before.view(np.uint32))
    assert st["candidate_pairs"] > 0


def test_initially_overlapping_voxels_are_projected_to_touching():
    # src/collision.rs:994-1015 (the projection inside that step): two voxels at one point
    s = line_box((2, 1, 1))
    s["bond_neighbor"][:] = EMPTY
    mid = (s["positions"][0] + s["positions"][1]) * np.float32(0.5)
    s["positions"][:] = mid
    s["previous"][:] = mid
    p, _, c, _ = oracle.contact_project(**s)
    assert p[1, 0] - p[0, 0] >= CELL - np.float32(1e-5)
    assert list(c["normal"][0]) == [1.0, 0.0, 0.0]


def brute_force_candidates(pos, prev, kin):
    """Pairs whose bin ranges (bin_bounds of the swept box) intersect and that
    are not both kinematic: the set hashed bins produce."""
    m = np.float32(0.5) + np.float32(2e-4)
    a = np.minimum(prev, pos) / CELL
    b = np.maximum(prev, pos) / CELL
    lo = np.floor(a - m).astype(np.int64)
    hi = np.floor(b + m).astype(np.int64)
    n = len(pos)
    out = []
    for i in range(n):
        ok = np.all((np.maximum(lo[i], lo[i + 1:]) <= np.minimum(hi[i], hi[i + 1:])), axis=1)
        ok &= ~(kin[i] & kin[i + 1:])
        out += [(i, i + 1 + int(j)) for j in np.nonzero(ok)[0]]
    return np.array(sorted(out), dtype=np.uint32).reshape(-1, 2)


@pytest.mark.parametrize("dims,seed", [((5, 4, 3), 1), ((6, 3, 4), 2), ((4, 4, 4), 3)])
def test_broad_phase_matches_brute_force(dims, seed):
    s = scenes.contact_box(dims, seed=seed, movers=0.2, mover_reach=2.5)
    got = oracle.contact_candidates(s["positions"], s["previous"], s["kinematic"])
    want = brute_force_candidates(s["positions"], s["previous"], s["kinematic"])
    assert np.array_equal(got, want)


def test_separation_conserves_momentum_and_resolves_overlap():
    s = scenes.contact_box((8, 7, 6), seed=4, fixed_floor=False, targeted=0.0, movers=0.0, jitter=6e-4)
    p, q, c, st = oracle.contact_project(**s)
    dp = p.astype(np.float64) - s["positions"]
    moved = np.abs(dp).sum()
    assert moved > 1e-4
    impulse = (s["mass"][:, None].astype(np.float64) * dp).sum(axis=0)
    assert np.abs(impulse).max() < 1e-6 * (s["mass"].astype(np.float64)[:, None] * np.abs(dp)).sum()
    # `separate` moves previous with positions: the velocity (p - q) is untouched
    assert np.allclose(p.astype(np.float64) - q, s["positions"].astype(np.float64) - s["previous"], atol=1e-9)
    # every pair that still overlaps after the last pass would have kept `changed` set
    if st["passes"] < 6:
        d = np.abs(p[c["i"]].astype(np.float64) - p[c["j"]])
        assert np.all(d.max(axis=1) >= 0.01 - 1e-6)


def test_kinematic_and_targeted_axes_do_not_move():
    s = scenes.contact_box((8, 6, 5), seed=5, targeted=0.3)
    p, q, _, _ = oracle.contact_project(**s)
    kin = s["kinematic"]
    assert np.array_equal(p[kin], s["positions"][kin])
    held = s["targets"]
    assert np.array_equal(p[held], s["positions"][held])
    assert not np.array_equal(p[~kin], s["positions"][~kin])


def test_intact_bond_pairs_are_never_projected():
    s = scenes.contact_box((7, 6, 5), seed=6, broken=0.05)
    _, _, c, _ = oracle.contact_project(**s)
    nb = s["bond_neighbor"]
    for a in range(3):
        bonded = {(int(i), int(nb[a, i])) for i in np.nonzero(nb[a] != EMPTY)[0]}
        assert not bonded & set(zip(c["i"].tolist(), c["j"].tolist()))
