"""GPU tests of the host-facing surface around the hot path: the server's
octet-stream wire formats, compiled programs with measure chains and the
cost-model plan choice, device expectation values, and noisy pure-state
evolution with stochastic Kraus maps and readout POVMs."""
import sys

import numpy as np
import pytest

from qvm_amd import _native as N
from qvm_amd import circuits as CI
from qvm_amd import gates as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def S():
    from qvm_amd.state import HipPureState
    assert N.device_count() >= 1, "no HIP device visible"
    return HipPureState


def _layered_state(S, oracle, n, dtype, seed=4):
    gates = CI.layered(n, 3, seed=seed)
    st = S(n, dtype)
    ref = oracle.zero_state(n)
    for U, q in gates:
        st.apply_matrix(U, G.nat_tuple(*q))
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    return st, ref


@pytest.mark.parametrize("n", [3, 14, 23])
def test_wire_formats_match_big_endian_encoding(S, oracle, n):
    """wavefunction and probabilities replies: big-endian IEEE doubles,
    byte-for-byte equal to encoding the oracle's amplitudes (n=23 spans
    several export chunks of the double-buffered copy)."""
    say(f"wire formats n={n}")
    st, ref = _layered_state(S, oracle, n, "f64")
    wf = st.wavefunction_octets()
    assert len(wf) == 16 * (1 << n)
    assert wf == ref.astype(">c16").tobytes()
    pr = st.probabilities_octets()
    p_ref = ref.real * ref.real + ref.imag * ref.imag  # probability, src/wavefunction.lisp:44-50
    assert pr == p_ref.astype(">f8").tobytes()
    # sub-range and native-order probabilities
    a, c = (1 << n) // 3, max(1, (1 << n) // 5)
    assert st.wavefunction_octets(a, c) == ref[a:a + c].astype(">c16").tobytes()
    np.testing.assert_array_equal(st.probabilities(a, c), p_ref[a:a + c])
    np.testing.assert_array_equal(st.read_amplitudes(a, c), ref[a:a + c])
    st.close()


def test_wire_formats_f32_widen_to_double(S, oracle):
    n = 12
    st, ref = _layered_state(S, oracle, n, "f32")
    amps = st.read_amplitudes()
    assert np.max(np.abs(amps - ref)) < 1e-5
    wf = np.frombuffer(st.wavefunction_octets(), dtype=">c16")
    np.testing.assert_array_equal(wf, amps)  # the widened FP32 values, exactly
    pr = np.frombuffer(st.probabilities_octets(), dtype=">f8")
    np.testing.assert_array_equal(pr, st.probabilities())
    st.close()


def test_compiled_measure_chain_is_one_measure_all(S):
    from qvm_amd import qvm as QV
    n = 10
    prog = "DECLARE ro BIT[10]\n" + "\n".join(f"H {k}" for k in range(n - 1)) + "\nCNOT 0 9\n" + \
        "\n".join(f"MEASURE {k} ro[{k}]" for k in range(n))
    draws = [0.3141592653589793]
    q = QV.make_qvm(n, rng=lambda: draws[0])
    q.load_program(prog)
    cp = q.compile_loaded_program(mode="tiled")
    assert any(isinstance(st, QV.MeasureAll) for st in cp.steps)
    q.run()
    assert q.draws_used == 1  # one draw for the whole chain
    bits = [int(q.memory_ref("ro", k)) for k in range(n)]
    # the state collapsed to the basis state of the recorded bits, and it is
    # the one the measure-all rule picks for the draw on the uniform-over-
    # {x : x0 == x9} distribution
    b = sum(bit << k for k, bit in enumerate(bits))
    assert q.nth_amplitude(b) == 1
    probs = np.array([(2.0 ** -9 if ((x & 1) == (x >> 9 & 1)) else 0.0) for x in range(1 << n)])
    cum = np.cumsum(probs)
    assert b == int(np.argmax(cum > draws[0]))
    q.close()
    # without the chain compilation every MEASURE draws (or skips at p1 == 0)
    q = QV.make_qvm(n, rng=lambda: draws[0])
    q.load_program(prog)
    q.compile_loaded_program(mode="plain", measure_chains=False)
    q.run()
    assert q.draws_used >= 2
    q.close()


def test_auto_plan_matches_plain(S, oracle):
    from qvm_amd import planner as P
    n = 16
    gates = CI.random_mix(n, 10, seed=11)
    st = S(n, "f64")
    say("auto plan n=16")
    cp = P.compile_gates(st, gates, mode="auto")
    cp.run()
    ref = oracle.zero_state(n)
    for U, q in gates:
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    assert np.max(np.abs(st.amplitudes() - ref)) < 1e-12
    cp.close()
    st.close()


def test_perform_expectation(S):
    from qvm_amd import qvm as QV
    prep = "RY(0.7) 0\nCNOT 0 1\nRX(1.1) 2\nH 3\nCNOT 3 1"
    ops = ["Z 0", "Z 0\nZ 1", "X 3", "Y 2", "I 0", "Z 2\nX 3"]
    vals = QV.perform_expectation(prep, ops, 4)
    # dense reference from the same gate matrices
    psi = np.zeros(16, dtype=np.complex128)
    psi[0] = 1

    def apply(psi, name, params, qs):
        U = G.gate_matrix(name, params)
        k = len(qs)
        # build the operator on 4 qubits by brute force over basis states
        out = np.zeros_like(psi)
        nat = G.nat_tuple(*qs)
        for x in range(16):
            col = sum(((x >> nat[j]) & 1) << j for j in range(k))
            for row in range(1 << k):
                y = x
                for j in range(k):
                    y = (y & ~(1 << nat[j])) | (((row >> j) & 1) << nat[j])
                out[y] += U[row, col] * psi[x]
        return out

    for line in prep.splitlines():
        nm, rest = (line.split(" ", 1) + [""])[:2]
        params = []
        if "(" in nm:
            nm, p = nm.split("(")
            params = [float(p.rstrip(")"))]
        psi = apply(psi, nm, params, [int(x) for x in rest.split()])
    for op, v in zip(ops, vals):
        phi = psi.copy()
        for line in op.splitlines():
            nm, qs = line.split(" ", 1)
            phi = apply(phi, nm, [], [int(x) for x in qs.split()])
        assert abs(v - np.vdot(psi, phi).real) < 1e-14, op


def test_stochastic_kraus_selection_and_state(S, oracle):
    """Pauli noise: p_j of K_j = sqrt(w_j) P_j is w_j whatever the state, so
    the chosen operator is min{j : w_1 + ... + w_j >= r}; the resulting state
    is P_j psi (unit norm).  Amplitude damping: p_1 = gamma * P(q = 1)."""
    from qvm_amd import noise as NZ
    n = 9
    rng = np.random.default_rng(5)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi = v / np.linalg.norm(v)
    ops = NZ.make_pauli_noise_map(0.1, 0.2, 0.3)
    w = np.array([0.4, 0.1, 0.2, 0.3])
    cw = np.cumsum(w)
    st = S(n, "f64")
    trial = S(n, "f64")
    for r in (0.05, 0.39, 0.41, 0.49, 0.51, 0.69, 0.71, 0.99):
        st.write_amplitudes(psi)
        j = NZ.evolve_pure_state_stochastically(st, trial, ops, [4], r)
        assert j == int(np.argmax(cw >= r)), r
        ref = psi.copy()
        oracle.apply_matrix(ref, ops[j], [4])
        ref = ref / np.linalg.norm(ref)
        assert np.max(np.abs(st.amplitudes() - ref)) < 1e-14
    gamma = 0.35
    K0 = np.array([[1, 0], [0, np.sqrt(1 - gamma)]], dtype=np.complex128)
    K1 = np.array([[0, np.sqrt(gamma)], [0, 0]], dtype=np.complex128)
    NZ.check_kraus_ops([K0, K1])
    p1 = float(np.sum(np.abs(psi.reshape(-1, 2, 8)[:, 1, :]) ** 2))  # P(qubit 3 = 1)
    p_k0 = 1 - gamma * p1
    for r, want in ((p_k0 - 1e-3, 0), (p_k0 + 1e-3, 1)):
        st.write_amplitudes(psi)
        j = NZ.evolve_pure_state_stochastically(st, trial, [K0, K1], [3], r)
        assert j == want
        ref = psi.copy()
        oracle.apply_matrix(ref, [K0, K1][j], [3])
        ref = ref / np.linalg.norm(ref)
        assert np.max(np.abs(st.amplitudes() - ref)) < 1e-14
    st.close()
    trial.close()


def test_noisy_qvm_gates_and_readout():
    from qvm_amd import noise as NZ
    # a deterministic draw sequence: Kraus draw, then the measurement draws
    seq = iter([0.95, 0.5, 0.2, 0.5, 0.97])
    q = NZ.NoisyQVM(2, rng=lambda: next(seq))
    q.set_noisy_gate("X", [0], NZ.make_pauli_perturbed_1q_gate("X", 0.0, 0.0, 0.1))
    q.set_readout_povm(1, (0.9, 0.2, 0.1, 0.8))
    q.load_program("DECLARE ro BIT[2]\nX 0\nMEASURE 0 ro[0]\nMEASURE 1 ro[1]")
    q.run()
    # r = 0.95 > 0.9 picks sqrt(0.1) Z X: |1> up to phase on qubit 0
    assert int(q.memory_ref("ro", 0)) == 1
    # (draw 0.5 <= p1 = 1); qubit 1 is |0>: p1 == 0 so its MEASURE draws
    # nothing, and the POVM draw 0.2 <= p(0|0) keeps the recorded 0
    assert int(q.memory_ref("ro", 1)) == 0
    _, bits = q.measure_all()  # draw 0.5 samples |01>; POVM draw 0.97 > p(0|0) flips qubit 1
    assert bits == [1, 1]
    with pytest.raises(Exception):
        q.set_readout_povm(0, (0.5, 0.5, 0.6, 0.5))
    with pytest.raises(Exception):
        q.set_noisy_gate("H", [1], [np.eye(2) * 0.5])
    q.close()
