"""Assembly of (M/h^2 + J^T K J) dx = -J^T K C (src/solver.rs:1539-1636) and
its row packing (src/solver.rs:1652-1672, 1731-1777), checked against finite
differences of the energy whose Hessian and negative gradient the system is:

  E(x) = sum_i m_i/(2h^2) |x_i - pred_i|^2 - sum_i load_i . x_i
       + sum_bonds 1/2 (x_j - x_i - rest)^T k (x_j - x_i - rest)

restricted to the free DOFs. A wrong sign, a transposed block, a swapped
endpoint or a missing projection changes the Hessian or the gradient."""
import numpy as np
import pytest

EMPTY = 0xFFFFFFFF


def perturbed_scene(scenes, seed):
    rng = np.random.default_rng(seed)
    s = scenes.mixed_box((3, 3, 3), seed=seed, fixed_face="x-")
    s.positions = (s.positions + rng.normal(scale=2e-4, size=s.positions.shape)).astype(np.float32)
    s.targets[13, 1] = s.positions[13, 1]
    s.targets[22] = s.positions[22]
    s.loads[5] = np.array([3.0, -2.0, 1.0], np.float32)
    s.loads[s.kinematic != 0] = 0
    s.damage[0, 4] = 0.5
    s.plastic_strain[1, 10] = np.array([1e-4, -2e-4, 5e-5], np.float32)
    s.formation_offset[2, 3] = np.array([-1e-5, 2e-5, 0.0], np.float32)
    return s


def energy_fn(sys_, s, h, pred):
    c = sys_.constraints()
    fm = s.free_mask()
    both_fixed = (fm[c["ija"][:, 0]] == 0) & (fm[c["ija"][:, 1]] == 0)
    keep = ~both_fixed
    i = c["ija"][keep, 0].astype(np.int64)
    j = c["ija"][keep, 1].astype(np.int64)
    k = c["k"][keep].reshape(-1, 3, 3)
    rest = c["rest"][keep]
    m = s.mass.astype(np.float64) / (h * h)
    loads = s.loads.astype(np.float64).copy()
    loads[s.kinematic != 0] = 0

    def E(x):
        x = x.reshape(-1, 3)
        e = x[j] - x[i] - rest
        return (0.5 * np.sum(m[:, None] * (x - pred) ** 2) - np.sum(loads * x)
                + 0.5 * np.einsum("bi,bij,bj->", e, k, e))
    return E


def fd_grad(E, x, d):
    g = np.zeros_like(x)
    for q in range(x.size):
        xp = x.copy()
        xm = x.copy()
        xp[q] += d
        xm[q] -= d
        g[q] = (E(xp) - E(xm)) / (2 * d)
    return g


@pytest.mark.parametrize("seed", [1, 2])
def test_assembly_is_hessian_and_gradient_of_energy(O, scenes, seed):
    s = perturbed_scene(scenes, seed)
    h = 1e-3
    os_ = O.OracleScene(s)
    O.fit_rotations(os_, O.F64)
    pred = os_.positions.copy()
    pred[:, 2] += 1e-5
    sys_ = O.assemble(os_, O.F64, h, pred, s.free_mask(), s.loads)
    E = energy_fn(sys_, s, h, pred)
    x0 = os_.positions.ravel().copy()
    n3 = x0.size
    d = 1e-6
    grad = fd_grad(E, x0, d)
    # Hessian columns from central differences of the (exact) gradient.
    H = np.zeros((n3, n3))
    for q in range(n3):
        xp = x0.copy()
        xm = x0.copy()
        xp[q] += d
        xm[q] -= d
        H[:, q] = (fd_grad(E, xp, d) - fd_grad(E, xm, d)) / (2 * d)
    fm = s.free_mask()
    free = np.array([[(f >> a) & 1 for a in range(3)] for f in fm], np.float64).ravel()
    full = sys_.full()
    rows = sys_.rows()
    scale = np.abs(H).max()
    # rhs = -P grad E(p)
    np.testing.assert_allclose(full["rhs"].ravel(), -free * grad, atol=1e-6 * np.abs(grad).max())
    # projected diagonal blocks: P H_ii P + I - P
    for v in range(s.n):
        P = np.diag(free[3 * v:3 * v + 3])
        Hii = H[3 * v:3 * v + 3, 3 * v:3 * v + 3]
        np.testing.assert_allclose(full["diagonal"][v].reshape(3, 3), P @ Hii @ P + np.eye(3) - P,
                                   atol=1e-6 * scale)
    # packed off-diagonal blocks: P_i H_ij P_j at each neighbour slot
    for slot, v in enumerate(rows["active"]):
        Pi = np.diag(free[3 * v:3 * v + 3])
        seen = set()
        for e in range(6):
            nb = rows["neighbors"][slot, e]
            blk = rows["blocks"][slot, 9 * e:9 * e + 9].reshape(3, 3)
            if nb == EMPTY:
                assert np.all(blk == 0)
                continue
            u = int(rows["active"][nb])
            seen.add(u)
            Pj = np.diag(free[3 * u:3 * u + 3])
            np.testing.assert_allclose(blk, Pi @ H[3 * v:3 * v + 3, 3 * u:3 * u + 3] @ Pj, atol=1e-6 * scale)
        # every active row coupled to v in the Hessian has a slot
        for u in rows["active"]:
            if int(u) == int(v) or int(u) in seen:
                continue
            Pj = np.diag(free[3 * u:3 * u + 3])
            assert np.abs(Pi @ H[3 * v:3 * v + 3, 3 * u:3 * u + 3] @ Pj).max() <= 1e-6 * scale


def test_f32_assembly_matches_f64(O, scenes):
    s = perturbed_scene(scenes, 3)
    out = {}
    for mode in (O.F32, O.F64):
        os_ = O.OracleScene(s)
        O.fit_rotations(os_, mode)
        out[mode] = O.assemble(os_, mode, 1e-3, os_.positions.copy(), s.free_mask(), s.loads).rows()
    assert np.array_equal(out[O.F32]["neighbors"], out[O.F64]["neighbors"])
    for key in ("blocks", "diagonal"):
        sc = np.abs(out[O.F64][key]).max()
        assert np.abs(out[O.F32][key] - out[O.F64][key]).max() <= 1e-5 * sc


def test_inverse_packing_is_inverse_of_diagonal(O, scenes):
    s = perturbed_scene(scenes, 4)
    os_ = O.OracleScene(s)
    rows = O.assemble(os_, O.F64, 1e-3, os_.positions.copy(), s.free_mask(), s.loads).rows()
    d = rows["diagonal"].reshape(-1, 3, 3)
    inv = rows["inverse"].reshape(-1, 3, 3)
    np.testing.assert_allclose(np.einsum("nij,njk->nik", d, inv), np.broadcast_to(np.eye(3), d.shape), atol=1e-9)
