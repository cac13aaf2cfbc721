"""Pins the CPU oracle (oracle/qvm_oracle.c) before the GPU results are checked against it.

Each check is chosen so that a plausible slip in the restatement fails: the
reference's own golden values, brute-force Kronecker-product unitaries with
asymmetric gates on non-adjacent, reversed-order qubits, closed forms, and
cross-mode agreement (serial / generic / compiled / AVX-FMA kernels,
serial vs threaded runtime).
"""
import json
import math
import os

import numpy as np
import pytest

from qvm_amd import circuits as C
from qvm_amd import gates as G
from qvm_amd import quil as Q

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def kron_unitary(U, quil_qubits, n):
    """Full 2^n unitary of U on quil_qubits by brute force over basis states
    (the program-matrix idea of src/misc.lisp:9-23)."""
    nat = G.nat_tuple(*quil_qubits)
    D = 1 << n
    M = np.zeros((D, D), dtype=np.complex128)
    k = len(nat)
    for col in range(D):
        gin = sum(((col >> q) & 1) << j for j, q in enumerate(nat))
        rest = col
        for q in nat:
            rest &= ~(1 << q)
        for r in range(1 << k):
            row = rest
            for j, q in enumerate(nat):
                if (r >> j) & 1:
                    row |= 1 << q
            M[row, col] += U[r, gin]
    return M


def random_unitary(k, rng):
    Z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    Qm, R = np.linalg.qr(Z)
    return Qm * (np.diag(R) / np.abs(np.diag(R)))


def random_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return np.ascontiguousarray(v / np.linalg.norm(v))


def run_quil_oracle(oracle, lines, n):
    prog = Q.parse_quil("\n".join(lines))
    psi = oracle.zero_state(n)
    for ins in prog.executable_code:
        U = Q.resolve_gate_matrix(ins, prog.gate_definitions, {})
        oracle.apply_matrix(psi, U, G.nat_tuple(*ins.qubits))
    return psi


# --- index math goldens ----------------------------------------------------

def test_inject_eject_index_goldens(oracle):
    g = GOLD["inject_bit"]
    for n, want in g["cases"]:
        assert oracle.inject_bit(g["x"], n) == want
    g = GOLD["eject_bit"]
    for x, n in g["cases"]:
        assert oracle.eject_bit(x, n) == g["target"]
    g = GOLD["index_to_address"]
    for q, want in g["cases"]:
        assert oracle.index_to_address(g["index"], q, g["state"]) == want


def test_nat_tuple_and_subdivide_goldens(oracle):
    g = GOLD["nat_tuple"]
    assert oracle.nat_tuple(*g["args"]) == g["expected"]
    assert list(G.nat_tuple(*g["args"])) == g["expected"]
    g = GOLD["subdivide"]
    for total, ranges in g["cases"]:
        assert oracle.subdivide(total, g["workers"]) == [tuple(r) for r in ranges]


def test_transposition_decomposition_round_trip(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        d = 1 << int(rng.integers(1, 6))
        perm = list(rng.permutation(d))
        ident = list(range(d))
        for a, b in oracle.permutation_to_transpositions(perm):
            ident[a], ident[b] = ident[b], ident[a]
        assert ident == perm


# --- sampling / CDF goldens ------------------------------------------------

def test_sample_bisect_goldens(oracle):
    for case in GOLD["sample_bisect"]["cases"]:
        psi = np.array(case["amplitudes"], dtype=np.complex128)
        for p, want in case["draws"]:
            assert oracle.sample_bisect(psi, p) == want, (case["probs"], p)


def test_cdf_goldens(oracle):
    for case in GOLD["cdf"]["cases"][1:]:
        psi = np.array(case["amps"], dtype=np.complex128)
        # cdf only needs a vector; oracle.cdf works on any length
        assert np.allclose(oracle.cdf(psi), case["cdf"], rtol=1e-12)


def test_cdf_search_tie_rule_matches_brute_force(oracle):
    rng = np.random.default_rng(1)
    for n in range(1, 9):
        psi = random_state(n, rng)
        psi[rng.integers(0, psi.size, size=psi.size // 3)] = 0  # zero-probability runs
        c = oracle.cdf(psi)
        draws = list(rng.random(64)) + list(c[:8]) + [0.0, 1.0, float(c[-1])]
        for p in draws:
            want = next((i for i in range(psi.size) if c[i] >= p), psi.size - 1)
            assert oracle.cdf_search(c, p) == want


def test_bisect_rule_matches_brute_force(oracle):
    rng = np.random.default_rng(2)
    for n in range(1, 9):
        psi = random_state(n, rng)
        psi[rng.integers(0, psi.size, size=psi.size // 3)] = 0
        pr = np.abs(psi) ** 2
        for p in list(rng.random(40)) + [0.0, 1.0]:
            # min b with C(b) > p under the bisection's own partial sums; brute
            # force with exact (long double) prefix is equal except at ties
            C_ = np.cumsum(pr.astype(np.longdouble))
            want = next((b for b in range(psi.size) if C_[b] > p), psi.size - 1)
            got = oracle.sample_bisect(psi, p)
            if got != want:
                # only legitimate when p sits on a boundary within rounding
                assert abs(float(C_[min(got, want)]) - p) < 1e-14


def test_matmul2_golden(oracle):
    g = GOLD["matmul2"]
    U = np.array([[complex(*z) for z in row] for row in g["matrix"]])
    psi = np.array([complex(*z) for z in g["vector"]])
    want = np.array([complex(*z) for z in g["expected"]])
    for fn in (oracle.apply_matrix, oracle.apply_compiled, oracle.apply_avx):
        out = fn(psi.copy(), U, [0])
        assert np.array_equal(out, want)


# --- gate application ------------------------------------------------------

@pytest.mark.parametrize("n", [3, 5, 7])
def test_gates_match_brute_force_unitary(oracle, n):
    rng = np.random.default_rng(10 + n)
    psi = random_state(n, rng)
    ref = psi.copy()
    for k in (1, 2, 3):
        for _ in range(4):
            qs = list(rng.choice(n, size=k, replace=False))
            U = random_unitary(k, rng)
            M = kron_unitary(U, qs, n)
            ref = M @ ref
            oracle.apply_matrix(psi, U, G.nat_tuple(*qs))
    assert np.max(np.abs(psi - ref)) < 1e-13


def test_cnot_orientation_is_asymmetric(oracle):
    # CNOT 2 0: control 2, target 0.  |100> -> |101>
    psi = oracle.zero_state(3)
    psi[:] = 0
    psi[0b100] = 1
    oracle.apply_matrix(psi, G.gate_matrix("CNOT"), G.nat_tuple(2, 0))
    assert psi[0b101] == 1 and abs(psi).sum() == 1


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_modes_agree(oracle, k):
    rng = np.random.default_rng(20 + k)
    n = 8
    for trial in range(3):
        qs = list(rng.choice(n, size=k, replace=False))
        U = random_unitary(k, rng)
        if trial == 1:
            U[np.abs(U) < 0.3] = 0  # zeros exercise the compiled mode's skipping
        psi = random_state(n, rng)
        a = oracle.apply_matrix(psi.copy(), U, qs)
        b = oracle.apply_compiled(psi.copy(), U, qs)
        assert np.array_equal(a, b)  # skipping 0 / eliding x1 is exact
        if k <= 2:
            c = oracle.apply_avx(psi.copy(), U, qs)
            assert np.max(np.abs(a - c)) < 1e-15


def test_threaded_runtime_matches_serial(oracle):
    # the reference's trick: force the parallel branch on a small input
    rng = np.random.default_rng(3)
    n = 10
    psi = random_state(n, rng)
    gates = [(random_unitary(1, rng), [int(q)]) for q in rng.integers(0, n, 6)]
    gates += [(random_unitary(2, rng), list(rng.choice(n, 2, replace=False))) for _ in range(4)]
    serial = psi.copy()
    for U, q in gates:
        oracle.apply_matrix(serial, U, q)
    p_serial = [oracle.excited_probability(serial, q) for q in range(n)]
    try:
        oracle.set_parallelization(1, 4)
        par = psi.copy()
        for U, q in gates:
            oracle.apply_matrix(par, U, q)
        comp = psi.copy()
        for U, q in gates:
            oracle.apply_compiled(comp, U, q)
        p_par = [oracle.excited_probability(par, q) for q in range(n)]
    finally:
        oracle.set_parallelization(19, 0)
    assert np.array_equal(serial, par)
    assert np.array_equal(serial, comp)
    assert np.allclose(p_serial, p_par, rtol=1e-14, atol=0)


def test_permutation_gate_matches_matrix(oracle):
    rng = np.random.default_rng(4)
    n = 6
    for k in (1, 2, 3):
        perm = list(rng.permutation(1 << k))
        qs = list(rng.choice(n, size=k, replace=False))
        U = np.zeros((1 << k, 1 << k), dtype=np.complex128)
        for j, p in enumerate(perm):
            U[j, p] = 1  # psi'[j] = psi[perm[j]]
        psi = random_state(n, rng)
        a = oracle.apply_permutation(psi.copy(), perm, qs)
        b = oracle.apply_matrix(psi.copy(), U, qs)
        assert np.array_equal(a, b)


# --- closed forms ------------------------------------------------------------

def test_hadamard_uniform(oracle):
    for n in range(1, 10):
        psi = run_quil_oracle(oracle, [f"H {q}" for q in range(n)], n)
        assert np.allclose(np.abs(psi) ** 2, 2.0 ** -n, rtol=1e-12)


def test_ghz_and_bell(oracle):
    for n in range(1, 11):
        psi = run_quil_oracle(oracle, ["H 0"] + [f"CNOT 0 {i}" for i in range(1, n)], n)
        assert abs(abs(psi[0]) ** 2 - 0.5) < 1e-15 and abs(abs(psi[-1]) ** 2 - 0.5) < 1e-15
    g = GOLD["bell_state"]
    psi = run_quil_oracle(oracle, ["H 0", "CNOT 0 1"], 2)
    assert abs(psi[0] - g["amp0"]) < 1e-15 and abs(psi[3] - g["amp3"]) < 1e-15


def test_qft2_golden_table(oracle):
    g = GOLD["qft2"]
    for t in range(4):
        prep = [f"X {q}" for q in range(2) if (t >> q) & 1]
        psi = oracle.zero_state(2)
        for line in prep:
            oracle.apply_matrix(psi, G.gate_matrix("X"), [int(line.split()[1])])
        for U, q in C.qft(2):
            oracle.apply_matrix(psi, U, G.nat_tuple(*q))
        want = np.array([complex(*z) for z in g["expected"][t]])
        assert np.max(np.abs(psi - want)) < 1e-15


def test_qft_closed_form(oracle):
    # QFT (with bit reversal) of |x> is the Fourier vector exp(2 pi i x y / N)/sqrt(N)
    n = 6
    N = 1 << n
    for x in (0, 1, 5, 37):
        psi = oracle.zero_state(n)
        psi[:] = 0
        psi[x] = 1
        for U, q in C.qft(n):
            oracle.apply_matrix(psi, U, G.nat_tuple(*q))
        want = np.exp(2j * np.pi * x * np.arange(N) / N) / math.sqrt(N)
        assert np.max(np.abs(psi - want)) < 1e-13


def test_rotation_and_inversion(oracle):
    psi = run_quil_oracle(oracle, ["RX(pi/2) 0"] * 4, 1)
    assert abs(abs(psi[0]) ** 2 - 1) < 1e-15
    psi = run_quil_oracle(oracle, ["X 0", "X 1", "X 2"], 3)
    assert psi[-1] == 1
    psi = run_quil_oracle(oracle, ["X 0", "H 1", "CZ 0 1", "H 1"], 2)
    assert abs(abs(psi[-1]) ** 2 - 1) < 1e-15


def test_dagger_round_trip_golden(oracle):
    psi = run_quil_oracle(oracle, GOLD["dagger_inversion"]["program"], 3)
    assert abs(psi[0] - 1) < 1e-14 and np.max(np.abs(psi[1:])) < 1e-14


def path_sum(lines, n):
    """Amplitudes of a program from |0...0> as a sum over computational-basis
    paths: each gate maps every basis state x with amplitude a to the states
    y that differ from x on the gate's qubits only, with weight U[y, x], the
    gate matrix indexed with the gate's first qubit as the most significant
    bit (Quil's convention).  No tensor products and no nat-tuple reversal,
    so it checks the engine's qubit ordering independently."""
    prog = Q.parse_quil("\n".join(lines))
    amps = {0: 1.0 + 0j}
    for ins in prog.executable_code:
        U = Q.resolve_gate_matrix(ins, prog.gate_definitions, {})
        qs = list(ins.qubits)
        k = len(qs)
        nxt = {}
        for x, a in amps.items():
            col = sum(((x >> q) & 1) << (k - 1 - j) for j, q in enumerate(qs))
            base = x
            for q in qs:
                base &= ~(1 << q)
            for row in range(1 << k):
                w = U[row, col]
                if w == 0:
                    continue
                y = base
                for j, q in enumerate(qs):
                    if (row >> (k - 1 - j)) & 1:
                        y |= 1 << q
                nxt[y] = nxt.get(y, 0) + w * a
        amps = nxt
    out = np.zeros(1 << n, dtype=np.complex128)
    for x, a in amps.items():
        out[x] = a
    return out


def test_path_sum_programs_oracle(oracle):
    """The reference's MADPI programs (tests/path-simulate-tests.lisp:21-30):
    the oracle's wavefunction equals the path sum."""
    for lines in GOLD["path_sum_programs"]["programs"]:
        n = 1 + max(q for ins in Q.parse_quil("\n".join(lines)).executable_code for q in ins.qubits)
        assert np.max(np.abs(run_quil_oracle(oracle, lines, n) - path_sum(lines, n))) < 1e-15, lines
    # a swapped control and target changes the result, so the check has teeth
    assert np.max(np.abs(path_sum(["H 0", "CNOT 0 1"], 2) - path_sum(["H 0", "CNOT 1 0"], 2))) > 0.5


def test_modifier_identities():
    X = G.gate_matrix("X")
    assert np.array_equal(G.controlled(X), G.gate_matrix("CNOT"))
    assert np.array_equal(G.controlled(G.controlled(X)), G.gate_matrix("CCNOT"))
    assert np.allclose(G.gate_matrix("XY", [math.pi / 2]), G.gate_matrix("PISWAP", [math.pi / 2]))
    assert np.allclose(G.gate_matrix("RX", [math.pi]), -1j * X)
    assert np.allclose(G.gate_matrix("H") @ G.gate_matrix("H"), np.eye(2))
    for name in G.STANDARD_GATES:
        npar = {"FSIM": 2, "U3": 3}.get(name, 1)
        try:
            U = G.gate_matrix(name)
        except ValueError:
            U = G.gate_matrix(name, [0.37 * (i + 1) for i in range(npar)])
        assert G.is_unitary(U), name


def test_qubit_ordering_program(oracle):
    lines = GOLD["qubit_ordering"]["program"]
    a = run_quil_oracle(oracle, lines, 4)
    M = np.eye(16, dtype=np.complex128)
    prog = Q.parse_quil("\n".join(lines))
    for ins in prog.executable_code:
        M = kron_unitary(Q.resolve_gate_matrix(ins, {}, {}), ins.qubits, 4) @ M
    e0 = np.zeros(16, dtype=np.complex128)
    e0[0] = 1
    assert np.max(np.abs(a - M @ e0)) < 1e-14


# --- reductions & measurement -----------------------------------------------

def test_reductions_vs_long_double(oracle):
    rng = np.random.default_rng(5)
    n = 12
    psi = random_state(n, rng)
    pr = (np.abs(psi.astype(np.clongdouble)) ** 2)
    idx = np.arange(psi.size)
    for q in (0, 5, 11):
        p1 = float(pr[(idx >> q) & 1 == 1].sum())
        p0 = float(pr[(idx >> q) & 1 == 0].sum())
        assert abs(oracle.excited_probability(psi, q) - p1) < 1e-15
        assert abs(oracle.ground_probability(psi, q) - p0) < 1e-15
    assert abs(oracle.norm(psi) - float(np.sqrt(pr.sum()))) < 1e-15
    v = psi * 3.0
    oracle.normalize(v)
    assert abs(np.linalg.norm(v) - 1) < 1e-14


def test_measure_rules(oracle):
    # p1 == 0: bit 0, no draw consumed (interpreted); compiled always draws
    psi = oracle.zero_state(2)
    bit, used = oracle.measure(psi, 1, 0.5, compiled=False)
    assert (bit, used) == (0, 0)
    psi = oracle.zero_state(2)
    bit, used = oracle.measure(psi, 1, 0.5, compiled=True)
    assert (bit, used) == (0, 1)
    # |+>: interpreted bit = (u <= p1); compiled keep-zero = (u < p0)
    for u, bi, bc in ((0.25, 1, 0), (0.75, 0, 1)):
        s = np.array([1, 1], dtype=np.complex128) / math.sqrt(2)
        bit, used = oracle.measure(s, 0, u, compiled=False)
        assert (bit, used) == (bi, 1)
        assert abs(abs(s[bit]) - 1) < 1e-15 and s[1 - bit] == 0
        s = np.array([1, 1], dtype=np.complex128) / math.sqrt(2)
        bit, _ = oracle.measure(s, 0, u, compiled=True)
        assert bit == bc
    # ties: interpreted u == p1 gives 1 (<=); compiled u == p0 gives 1 (not <)
    s0 = np.array([0.6, 0.8], dtype=np.complex128)
    p1, p0 = 0.8 * 0.8, 0.6 * 0.6
    assert oracle.measure(s0.copy(), 0, p1, compiled=False)[0] == 1
    assert oracle.measure(s0.copy(), 0, p0, compiled=True)[0] == 1
    assert oracle.measure(s0.copy(), 0, np.nextafter(p1, 2), compiled=False)[0] == 0
    assert oracle.measure(s0.copy(), 0, np.nextafter(p0, -1), compiled=True)[0] == 0


def test_measure_all_basis_golden(oracle):
    g = GOLD["measure_all_basis"]
    for i, lines in enumerate(g["programs"]):
        psi = run_quil_oracle(oracle, lines or ["I 0"], g["num_qubits"])
        b = oracle.measure_all(psi, 0.5)
        assert b == i and psi[i] == 1 and np.count_nonzero(psi) == 1


def test_sample_as_distribution_restated_as_written(oracle):
    # the reference's single-pass sampler restarts its draw index for every
    # amplitude, so every draw below the total ends at the last index
    psi = np.array([0.5, 0.5, 0.5, 0.5], dtype=np.complex128)
    out = oracle.sample_as_distribution(psi, [0.1, 0.3, 0.6, 0.9])
    assert list(out) == [3, 3, 3, 3]
