"""The C ABI driven in the call order of lisp/qvm-hip.lisp, through ctypes,
for one program per gate class the reference dispatches on
(src/apply-gate.lisp:109-181), against the oracle.

What each Lisp method does, restated here call for call:
* apply-gate-to-state (every gate class): the gate's matrix (cl-quil's
  gate-matrix: CONTROLLED U = diag(I, U) with the control as the first Quil
  argument, FORKED U(a, b) = diag(U(a), U(b)), DAGGER U = U^dagger, a sequence
  gate = the product of its gates in application order, a permutation gate =
  its permutation matrix) as a row-major quantum-operator, qvmhip_apply_matrix
  on the nat-tuple (the Quil arguments reversed);
* measure (the reference's base-qvm method around the binding's two
  methods): qvmhip_prob_excited, a draw only when p1 /= 0, u <= p1 -> 1,
  qvmhip_force_measure;
* measure-all-state: one draw, qvmhip_measure_all;
* the compiled projector: one draw, qvmhip_measure with the compiled rule.
The FP64 amplitudes after the gates must equal the oracle's bit for bit
(after a collapse, to 1e-15: the two sum p1 in different orders), every
measured bit must equal the oracle's under the same draws, and both sides
must consume the same number of draws."""
import numpy as np
import pytest

from qvm_amd import circuits as CI
from qvm_amd import gates as G

pytestmark = pytest.mark.gpu


class Draws:
    """The uniform stream (qvm's RANDOM), consumed in order."""

    def __init__(self, seed, n=64):
        self.u = CI.uniforms(seed, n)
        self.i = 0

    def next(self):
        x = float(self.u[self.i])
        self.i += 1
        return x


def perm_matrix(perm):
    k = len(perm)
    P = np.zeros((k, k), dtype=np.complex128)
    for j, p in enumerate(perm):
        P[j, p] = 1  # psi'[j] = psi[perm[j]]
    return P


def programs():
    rx = lambda t: G.gate_matrix("RX", [t])  # noqa: E731
    return {
        "simple-gate": [(G.gate_matrix("H"), [0]), (G.gate_matrix("CNOT"), [0, 3])],
        "parameterized-gate": [(G.gate_matrix("H"), [1]), (rx(0.37), [2]), (G.gate_matrix("CPHASE", [0.9]), [2, 1])],
        "sequence-gate": [(G.gate_matrix("H"), [0]), (G.gate_matrix("CNOT") @ np.kron(rx(0.2), G.gate_matrix("H")),
                                                      [3, 1])],
        "permutation-gate": [(G.gate_matrix("H"), [2]), (perm_matrix([0, 2, 3, 1]), [2, 4])],
        "controlled-gate": [(G.gate_matrix("H"), [0]), (G.controlled(rx(1.1)), [0, 4]),
                            (G.controlled(G.controlled(G.gate_matrix("X"))), [0, 4, 1])],
        "forked-gate": [(G.gate_matrix("H"), [3]), (G.forked(rx(0.3), rx(2.1)), [3, 0])],
        "dagger-gate": [(G.gate_matrix("H"), [1]), (G.dagger(G.gate_matrix("U3", [0.4, 1.2, -0.7])), [1]),
                        (G.dagger(G.gate_matrix("XY", [0.6])), [1, 2])],
        "compiled-matrix-gate-application": [(G.gate_matrix("H"), [4]), (G.gate_matrix("RZZ", [0.8]), [4, 2])],
        "compiled-permutation-gate-application": [(G.gate_matrix("H"), [0]), (G.gate_matrix("CSWAP"), [0, 3, 1])],
    }


@pytest.mark.parametrize("cls", list(programs()))
def test_lisp_call_order_per_gate_class(oracle, cls):
    from qvm_amd.state import HipPureState
    n = 5
    st = HipPureState(n, "f64")  # make-hip-pure-state: qvmhip_create
    ref = oracle.zero_state(n)
    for U, quil_args in programs()[cls]:
        nat = G.nat_tuple(*quil_args)
        st.apply_matrix(U, list(nat))  # apply-operator-on-device
        oracle.apply_matrix(ref, U, list(nat))
    np.testing.assert_array_equal(st.amplitudes(), ref)  # state-elements
    dr_engine, dr_oracle = Draws(11), Draws(11)
    # MEASURE q (interpreted): prob, conditional draw, force
    for q in (0, 3):
        p1 = st.prob_excited(q)
        assert abs(p1 - oracle.excited_probability(ref, q)) < 1e-15
        bit = 0 if p1 == 0 else int(dr_engine.next() <= p1)
        st.force_measure(q, bit, p1)
        obit, consumed = oracle.measure(ref, q, dr_oracle.u[dr_oracle.i])
        dr_oracle.i += consumed
        assert bit == obit, (cls, q)
    # the engine's p1 is a device reduction, the oracle's a sequential sum:
    # the collapse scale 1/sqrt(p) may differ in its last bit
    assert np.max(np.abs(st.amplitudes() - ref)) < 1e-15
    # the compiled projector: always one draw
    for q in (1, 2):
        u = dr_engine.next()
        bit, consumed = st.measure(q, u, compiled=True)
        assert consumed == 1
        obit, _ = oracle.measure(ref, q, u, compiled=True)
        dr_oracle.i += 1
        assert bit == obit
    assert dr_engine.i == dr_oracle.i
    assert np.max(np.abs(st.amplitudes() - ref)) < 1e-15
    # MEASURE-ALL: one draw, collapse
    u = dr_engine.next()
    b = st.measure_all(u)
    assert b == oracle.measure_all(ref, u)
    amps = st.amplitudes()
    assert amps[b] == 1 and np.count_nonzero(amps) == 1
    # RESET (set-to-zero-state)
    st.reset_zero()
    np.testing.assert_array_equal(st.amplitudes(), oracle.zero_state(n))
    st.close()
