"""Writes reference_known_answers.json: the inputs and expected outputs that
the reference's own device tests and probe hold for the PCG path, restated as
data (nothing from the reference is imported or run):

* two_node_pcg      validation/cuda_sparse_probe.py:268-343
* dot_600           src/cuda_solver.rs:518-553 (alpha 0.75 check included)
* ratio             validation/cuda_sparse_probe.py:192-211
* partial_sum_17    validation/cuda_sparse_probe.py:232-245
* synthetic_matvec  validation/cuda_sparse_probe.py:35-73 (tolerance 2e-3)
* thermal_rest      src/solver.rs:1113-1114, 1156-1160
* wood_limits       src/solver.rs:1046-1051
* tension_fixture   validation/mechanical-tests.txt:3 (loose regression)
"""
import json
import os

import numpy as np

EMPTY = 0xFFFFFFFF


def main():
    rhs = [1.0, 2.0, 3.0, 4.0, -1.0, 0.5]
    blocks = [0.0] * 108
    for node, edge in ((0, 0), (1, 1)):
        for axis in range(3):
            blocks[node * 54 + edge * 9 + axis * 3 + axis] = -0.2
    expected = [0.0] * 6
    for c in range(3):
        b0, b1 = rhs[c], rhs[3 + c]
        expected[c] = (2.0 * b0 + 0.2 * b1) / 3.96
        expected[3 + c] = (0.2 * b0 + 2.0 * b1) / 3.96
    i = np.arange(600 * 3)
    left = ((i % 97 + 1) * 0.03125).astype(np.float32).astype(np.float64)
    right = ((i % 53 + 1) * 0.015625).astype(np.float32).astype(np.float64)
    data = {
        "two_node_pcg": {
            "rhs": rhs,
            "neighbors": [1, EMPTY, EMPTY, EMPTY, EMPTY, EMPTY, EMPTY, 0, EMPTY, EMPTY, EMPTY, EMPTY],
            "diagonal": [2.0, 0, 0, 0, 2.0, 0, 0, 0, 2.0] * 2,
            "inverse": [0.5, 0, 0, 0, 0.5, 0, 0, 0, 0.5] * 2,
            "blocks": blocks,
            "expected": expected,
            "tolerance": 1e-5,
        },
        "dot_600": {"rows": 600, "expected": float(np.dot(left, right)), "alpha": 0.75, "rel_tol": 1e-12},
        "ratio": {"num": 4.5, "den": 1.5, "expected": 3.0},
        "partial_sum_17": {"values": list(range(1, 18)), "expected": 153.0},
        "synthetic_matvec": {
            "dims": [20, 25, 20],
            "diagonal_block": [2.5, 0.1, -0.05, 0.1, 3.0, 0.08, -0.05, 0.08, 4.0],
            "block": [700.0, 31.0, -12.0, 31.0, 520.0, 8.0, -12.0, 8.0, 430.0],
            "offsets": [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]],
            "x": "((i % 19) - 9) * 0.03125",
            "abs_tol": 2e-3,
        },
        "thermal_rest": {"iron_1e-5_100K": 1.001, "wood_transverse_100K": 1.003, "wood_parallel_100K": 1.0003,
                         "tol": 1e-6},
        "wood_limits": {"tension_along": 100e6, "tension_across": 2.5e6, "compression_along": 50e6,
                        "compression_across": 2e6, "shear_along": 12e6, "shear_across": 3e6},
        # validation/mechanical-tests.txt:3 (and validation/STATUS.md:13, the latest checkpoint)
        "tension_fixture": {"parallel_N": 1022767.250, "perpendicular_N": 25920.164, "ratio": 39.458,
                            "first_breaks": [900, 900]},
        # validation/STATUS.md:14, the latest checkpoint (validation/mechanical-tests.txt:4, ratio 32.539 and
        # first breaks 3/4, is the earlier checkpoint of validation/STATUS.md:38)
        "compression_fixture": {"ratio": 25.877},
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_known_answers.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
