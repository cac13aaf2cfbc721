"""Writes tests/golden/reference_vectors.json: the expected values that the
reference's own test files state for the pure-state path, transcribed as data
with their file:line.  The reference (Common Lisp) cannot run in this image,
so the values are taken from the assertions themselves; each entry's
"source" names the assertion it comes from.

Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def f32sqrt(xs):
    # (sqrt 0.1) in the test is a single-float sqrt of a single-float literal,
    # stored into a (complex double-float) vector.
    return [float(np.sqrt(np.float32(x))) for x in xs]


def main():
    g = {}
    g["inject_bit"] = {
        "source": "tests/utilities-tests.lisp:91-97",
        "x": 0b1111, "cases": [[0, 0b11110], [1, 0b11101], [2, 0b11011], [3, 0b10111], [4, 0b01111]]}
    g["eject_bit"] = {
        "source": "tests/utilities-tests.lisp:99-105",
        "target": 0b1111, "cases": [[0b11110, 0], [0b11101, 1], [0b11011, 2], [0b10111, 3], [0b01111, 4]]}
    g["index_to_address"] = {
        "source": "tests/measurement-tests.lisp:10-20",
        "index": 0b1111, "state": 0, "cases": [[0, 0b11110], [1, 0b11101], [2, 0b11011], [3, 0b10111], [4, 0b01111]]}
    g["nat_tuple"] = {"source": "tests/utilities-tests.lisp:25-28", "args": [1, 2, 3], "expected": [3, 2, 1]}
    W = 8
    slack = [[i, i + 1] for i in range(W - 1)] + [[W - 1, W + 3]]
    g["subdivide"] = {
        "source": "tests/utilities-tests.lisp:59-81", "workers": W,
        "cases": ([[i, [[0, i]]] for i in range(W)]
                  + [[W, [[i, i + 1] for i in range(W)]]]
                  + [[2 * W, [[i, i + 2] for i in range(0, 2 * W, 2)]]]
                  + [[W + 3, slack]])}
    g["sample_bisect"] = {
        "source": "tests/measurement-tests.lisp:184-211",
        "cases": [
            {"probs": [0.1] * 8,
             "draws": [[i * 0.1, i] for i in range(8)] + [[0.11, 1], [0.15, 1], [1.0, 7]]},
            {"probs": [0.0, 0.0, 0.0, 0.1, 0.5, 0.2, 0.1, 0.1],
             "draws": [[0.0, 3], [0.09999, 3], [0.1, 4], [0.10001, 4]]},
            {"probs": [0.2, 0.8], "draws": [[0.0, 0], [0.1, 0], [0.2, 1], [0.8, 1]]},
        ],
        "amplitude_rule": "float32 sqrt of each prob, widened to float64"}
    # the loop steps p by (+ p 0.1d0): recompute the exact doubles of that loop
    ps, p = [], 0.0
    for i in range(8):
        ps.append(p)
        p = p + 0.1
    g["sample_bisect"]["cases"][0]["draws"][:8] = [[ps[i], i] for i in range(8)]
    for c in g["sample_bisect"]["cases"]:
        c["amplitudes"] = f32sqrt(c["probs"])
    g["cdf"] = {"source": "tests/linear-algebra-tests.lisp:66-75",
                "cases": [{"amps": [], "cdf": []},
                          {"amps": [1, 1, 1], "cdf": [1.0, 2.0, 3.0]},
                          {"amps": [0.5 ** 0.5] * 3, "cdf": [0.5, 1.0, 1.5]}]}
    g["matmul2"] = {"source": "tests/qvm-avx-intrinsics.lisp:7-23",
                    "matrix": [[[1, 2], [3, 4]], [[5, 6], [7, 8]]], "vector": [[1, 2], [3, 4]],
                    "expected": [[-10, 28], [-18, 68]]}
    g["qft2"] = {"source": "tests/gate-tests.lisp:163-183",
                 "expected": [[[0.5, 0], [0.5, 0], [0.5, 0], [0.5, 0]],
                              [[0.5, 0], [0, 0.5], [-0.5, 0], [0, -0.5]],
                              [[0.5, 0], [-0.5, 0], [0.5, 0], [-0.5, 0]],
                              [[0.5, 0], [0, -0.5], [-0.5, 0], [0, 0.5]]],
                 "tolerance": 1e-4}
    g["measure_all_basis"] = {"source": "tests/measurement-tests.lisp:213-236", "num_qubits": 3,
                              "programs": [[("X %d" % q) for q in range(3) if (i >> q) & 1] for i in range(8)]}
    g["simple_measurements"] = {"source": "tests/measurement-tests.lisp:22-35",
                                "program": ["DECLARE ro BIT[3]", "X 0", "X 2", "MEASURE 0 ro[0]",
                                            "MEASURE 1 ro[1]", "MEASURE 2 ro[2]"], "ro": [1, 0, 1]}
    g["measurements_with_swap"] = {"source": "tests/measurement-tests.lisp:52-67",
                                   "program": ["DECLARE ro BIT[3]", "X 0", "X 2", "SWAP 0 1", "SWAP 0 2",
                                               "MEASURE 0 ro[0]", "MEASURE 1 ro[1]", "MEASURE 2 ro[2]"],
                                   "ro": [1, 1, 0]}
    g["bell_state"] = {"source": "tests/state-representation-tests.lisp:30-39",
                       "amp0": 0.5 ** 0.5, "amp3": 0.5 ** 0.5}
    g["dagger_inversion"] = {"source": "tests/modifier-tests.lisp:42-61",
                             "program": ["H 0", "CONTROLLED RY(2*pi/3) 0 1", "RX(pi/3) 0", "T 2", "CSWAP 0 1 2",
                                         "DAGGER CSWAP 0 1 2", "DAGGER T 2", "DAGGER RX(pi/3) 0",
                                         "DAGGER CONTROLLED RY(2*pi/3) 0 1", "DAGGER H 0"]}
    g["qubit_ordering"] = {"source": "tests/gate-tests.lisp:28-44",
                           "program": ["H 0", "H 1", "H 2", "H 3", "CNOT 2 0", "CSWAP 1 3 2"]}
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(g, f, indent=1)


if __name__ == "__main__":
    main()
