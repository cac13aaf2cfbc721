"""Noisy pure-state evolution: stochastic Kraus maps and classical readout
noise, on the HBM-resident engine.

Mirrors the reference's noisy-qvm (src/noisy-qvm.lisp) for pure states:

    make_pauli_noise_map          src/noisy-qvm.lisp:46-67
    make_pauli_perturbed_1q_gate  src/noisy-qvm.lisp:69-78
    NoisyQVM.set_noisy_gate       src/noisy-qvm.lisp:80-84
    NoisyQVM.set_readout_povm     src/noisy-qvm.lisp:86-89
    evolve_pure_state_stochastically  src/apply-gate.lisp:16-39
    check_kraus_ops / check_povm  src/channel-qvm.lisp:109-135, 197-206
    perturb_measurement           src/channel-qvm.lisp:185-195

The Kraus choice needs a second statevector of the same size (the trial
amplitudes): each candidate K_j is applied to a device copy of the state and
the running sum of its total probability is compared with one uniform draw;
the chosen trial buffer is swapped in and renormalised.  Every step is an
engine kernel (copy, gate application, reduction, scale).
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import gates as G
from . import quil as Q
from .qvm import PureStateQVM, QVMError, apply_gate_to_state
from .state import HipPureState


def check_kraus_ops(kraus_ops: Sequence[np.ndarray]) -> None:
    """Square, equal dimensions, sum_j K_j^H K_j = I within 1e-5."""
    ops = [np.asarray(k, dtype=np.complex128) for k in kraus_ops]
    if not ops:
        raise QVMError("empty Kraus map")
    m, n = ops[0].shape
    if m != n:
        raise QVMError("The Kraus operators be square matrices.")
    acc = np.zeros((m, n), dtype=np.complex128)
    for k in ops:
        if k.shape != (m, m):
            raise QVMError("All Kraus operators must have matching dimensions")
        acc += k.conj().T @ k
    if not np.allclose(acc, np.eye(m), atol=1e-5, rtol=0):
        raise QVMError("The Kraus map must preserve trace")


def check_povm(povm: Sequence[float]) -> None:
    """(p00 p01 p10 p11), each in [0, 1], columns summing to 1."""
    if len(povm) != 4:
        raise QVMError("a readout POVM is the list (p(0|0) p(0|1) p(1|0) p(1|1))")
    p00, p01, p10, p11 = (float(x) for x in povm)
    for p in (p00, p01, p10, p11):
        if not (0.0 <= p <= 1.0):
            raise QVMError("POVM entries must lie in [0, 1]")
    if abs(1.0 - (p00 + p10)) > 1e-10 or abs(1.0 - (p01 + p11)) > 1e-10:
        raise QVMError("POVM columns must sum to 1")


def make_pauli_noise_map(px: float, py: float, pz: float) -> List[np.ndarray]:
    """Kraus ops of rho -> (1-px-py-pz) rho + px X rho X + py Y rho Y + pz Z rho Z:
    sqrt(px) X, sqrt(py) Y, sqrt(pz) Z, then sqrt(1 - sum) I in front when the
    sum is below 1."""
    for p in (px, py, pz):
        if not (0.0 <= p <= 1.0):
            raise QVMError("Pauli probabilities must lie in [0, 1]")
    psum = px + py + pz
    if not (0.0 <= psum <= 1.0):
        raise QVMError("Pauli probabilities must sum to at most 1")
    ops = [np.sqrt(p) * G.gate_matrix(name) for name, p in (("X", px), ("Y", py), ("Z", pz))]
    if psum < 1:
        ops.insert(0, np.sqrt(1.0 - psum) * np.eye(2, dtype=np.complex128))
    return ops


def make_pauli_perturbed_1q_gate(gate_name: str, px: float, py: float, pz: float) -> List[np.ndarray]:
    """The ideal gate U followed by the Pauli noise map: K_j U for each K_j."""
    U = G.gate_matrix(gate_name)
    return [K @ U for K in make_pauli_noise_map(px, py, pz)]


def perturb_measurement(actual: int, p00: float, p01: float, p10: float, p11: float, r: float) -> int:
    """Recorded bit for true outcome ``actual`` and draw r: 0 iff r <= p(0|actual)."""
    if actual == 0:
        return 0 if r <= p00 else 1
    if actual == 1:
        return 0 if r <= p01 else 1
    raise QVMError("outcome must be a bit")


def evolve_pure_state_stochastically(state: HipPureState, trial: HipPureState, kraus_ops, quil_qubits,
                                     r: float) -> int:
    """Apply one Kraus operator of the map, chosen by inverse-transform
    sampling with the draw r: for j = 1, 2, ... the trial buffer is set to
    K_j psi and p_j = |K_j psi|^2 added to a running sum until the sum
    reaches r; the last trial buffer becomes the state, renormalised.
    Returns the index j chosen."""
    summed = 0.0
    chosen = -1
    for j, K in enumerate(kraus_ops):
        trial.copy_from(state)
        apply_gate_to_state(trial, np.asarray(K, dtype=np.complex128), quil_qubits)
        summed += trial.total_probability()
        chosen = j
        if summed >= r:
            break
    state.swap_contents(trial)
    state.normalize()
    return chosen


class NoisyQVM(PureStateQVM):
    """noisy-qvm (src/noisy-qvm.lisp:7-26): a pure-state QVM whose gates may
    have noisy Kraus realisations and whose qubits may have readout POVMs.
    Programs are not compiled (src/noisy-qvm.lisp:145-152)."""

    def __init__(self, num_qubits: int, noisy_gate_definitions: Dict | None = None,
                 readout_povms: Dict | None = None, **kw):
        super().__init__(num_qubits, **kw)
        self.noisy_gate_definitions: Dict[Tuple[str, Tuple[int, ...]], List[np.ndarray]] = {}
        self.readout_povms: Dict[int, Tuple[float, float, float, float]] = {}
        self._trial: HipPureState | None = None
        for (name, qubits), ops in (noisy_gate_definitions or {}).items():
            self.set_noisy_gate(name, qubits, ops)
        for q, povm in (readout_povms or {}).items():
            self.set_readout_povm(q, povm)

    def _trial_state(self) -> HipPureState:
        # check-allocate-computation-space: the trial buffer is made on first need
        if self._trial is None:
            self._trial = HipPureState(self.number_of_qubits, dtype=self.state.dtype)
        return self._trial

    def set_noisy_gate(self, gate_name: str, qubits, kraus_ops) -> None:
        ops = [np.asarray(k, dtype=np.complex128) for k in kraus_ops]
        check_kraus_ops(ops)
        self._trial_state()
        self.noisy_gate_definitions[(gate_name.upper(), tuple(int(q) for q in qubits))] = ops

    def set_readout_povm(self, qubit: int, povm) -> None:
        check_povm(povm)
        self.readout_povms[int(qubit)] = tuple(float(x) for x in povm)

    def transition(self, ins):
        if isinstance(ins, Q.GateApplication):
            key = (ins.name.upper(), tuple(ins.qubits))
            ops = self.noisy_gate_definitions.get(key)
            if ops is not None:
                if ins.modifiers:
                    raise QVMError("The noisy QVM doesn't support gate modifiers.")
                if ins.params:
                    raise QVMError("noisy gate definitions apply to static gates only")
                if ops[0].shape[0] != 1 << len(ins.qubits):
                    raise QVMError(f"Kraus operators of {ins.name} do not match its qubit count")
                evolve_pure_state_stochastically(self.state, self._trial_state(), ops, ins.qubits, self._draw())
                return self.pc + 1
            return super().transition(ins)
        if isinstance(ins, Q.Measure):
            nxt = super().transition(ins)
            # classical readout noise on the stored bit (src/channel-qvm.lisp:173-183)
            if ins.target is not None and ins.qubit in self.readout_povms:
                name, idx = ins.target
                c = int(self.memory[name][idx])
                self.memory[name][idx] = perturb_measurement(c, *self.readout_povms[ins.qubit], self._draw())
            return nxt
        return super().transition(ins)

    def measure_all(self):
        """measure-all with each bit passed through its POVM
        (perturb-measured-bits, src/measurement.lisp:228-243)."""
        qvm, bits = super().measure_all()
        out = []
        for i, c in enumerate(bits):
            povm = self.readout_povms.get(i)
            out.append(perturb_measurement(c, *povm, self._draw()) if povm else c)
        return qvm, out

    def compile_loaded_program(self, *a, **kw):
        return None

    def close(self) -> None:
        if self._trial is not None:
            self._trial.close()
            self._trial = None
        super().close()
