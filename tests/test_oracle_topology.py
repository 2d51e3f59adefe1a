"""Bond topology and block indexing of the oracle against the lattice rule
(src/scene.rs:558-570, 1668-1681) and the incidence-order slot rule
(src/solver.rs:1595-1630, 1652-1672), checked by brute force."""
import numpy as np
import pytest

EMPTY = 0xFFFFFFFF


def test_box_bond_counts(O, scenes):
    for dims, expected in [((10, 10, 10), 2700), ((64, 64, 64), 774_144), ((3, 2, 2), 20)]:
        s = scenes.uniform_box(dims)
        assert s.bond_count() == expected
        nb, count = O.build_topology(s.keys, s.kinematic)
        assert count == expected
        assert np.array_equal(nb, s.neighbor)


def test_spawn_kinematic_groups_are_not_bonded(O):
    # A 2x2x2 dynamic block on a 2x2x1 ground spawned fixed (the reference's
    # solver test scene): no bond crosses between the groups.
    keys = [(x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    keys += [(x, y, -1) for y in range(2) for x in range(2)]
    kin = np.array([0] * 8 + [1] * 4, np.uint8)
    nb, count = O.build_topology(np.array(keys), kin)
    assert count == 12 + 4  # 12 inside the block, 4 inside the ground slab
    for a in range(3):
        for i in range(12):
            j = nb[a, i]
            if j != EMPTY:
                assert kin[i] == kin[j]
    # Same cells with equal flags: the vertical bonds between the groups exist.
    nb2, count2 = O.build_topology(np.array(keys), np.zeros(12, np.uint8))
    assert count2 == count + 4


def test_arbitrary_spawn_order(O):
    rng = np.random.default_rng(0)
    keys = np.array([(x, y, z) for x in range(3) for y in range(4) for z in range(2)])
    perm = rng.permutation(len(keys))
    keys = keys[perm]
    kin = np.zeros(len(keys), np.uint8)
    nb, _ = O.build_topology(keys, kin)
    where = {tuple(k): i for i, k in enumerate(keys)}
    for i, k in enumerate(keys):
        for a in range(3):
            kk = list(k)
            kk[a] += 1
            assert nb[a, i] == where.get(tuple(kk), EMPTY)


def _brute_slots(s, free_mask):
    """Row slots by a direction table: for each active voxel, the bonds at
    -x, +x, -y, +y, -z, +z in that order (absent bonds compacted out, bonds
    whose both endpoints are fully fixed dropped), the neighbour's active
    index or EMPTY."""
    n = s.n
    active = np.nonzero(free_mask)[0]
    aidx = np.full(n, EMPTY, np.int64)
    aidx[active] = np.arange(active.size)
    incoming = {(a, int(j)): i for a in range(3) for i in range(n) for j in [s.neighbor[a, i]] if j != EMPTY}
    rows = np.full((active.size, 6), EMPTY, np.uint32)
    for slot, v in enumerate(active):
        e = 0
        for a in range(3):
            for other in (incoming.get((a, int(v))), None if s.neighbor[a, v] == EMPTY else int(s.neighbor[a, v])):
                if other is None:
                    continue
                if free_mask[v] == 0 and free_mask[other] == 0:
                    continue
                rows[slot, e] = aidx[other]
                e += 1
    return active, rows


@pytest.mark.parametrize("mode", [0, 1])
def test_incidence_order_matches_direction_table(O, scenes, mode):
    s = scenes.mixed_box((5, 4, 3), seed=11, fixed_face="z-")
    # One axis-targeted voxel and a fully targeted one (free mask 0, not kinematic).
    s.targets[7, 0] = s.positions[7, 0]
    s.targets[30] = s.positions[30]
    fm = s.free_mask()
    os_ = O.OracleScene(s)
    sy = O.assemble(os_, mode, 1e-3, os_.positions, fm, s.loads)
    rows = sy.rows()
    active, expected = _brute_slots(s, fm)
    assert np.array_equal(rows["active"], active)
    assert np.array_equal(rows["neighbors"], expected)
    # Empty slots carry zero blocks; inactive neighbours carry zero blocks.
    blocks = rows["blocks"].reshape(-1, 6, 9)
    assert np.all(blocks[rows["neighbors"] == EMPTY] == 0.0)
