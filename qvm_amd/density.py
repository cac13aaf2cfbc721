"""Density-matrix machine on the HBM-resident engine.

Mirrors the reference's density-qvm (src/density-qvm.lisp) and the density
methods of the state protocol: an n-qubit density matrix rho is held as
vec(rho), row-major (rho[i][j] at index i * 2^n + j), in a 2n-qubit engine
state, so every operation is one of the pure-state kernels:

    gate U on qubits q          U* on q, then U on the ghosts q + n
                                (vec(A rho B) = (A (x) B^T) vec(rho),
                                src/apply-gate.lisp:40-65)
    Kraus list {K_j}            sum_j (K_j (x) K_j*) vec(rho): a pristine copy,
                                a device running sum (qvmhip_accumulate)
                                (src/apply-gate.lisp:66-99)
    measurement probabilities   the diagonal Re rho[i][i] (qvmhip_density_diagonal,
                                src/state-representation.lisp:268-286)
    force-measurement           zero rows and columns of the annihilated
                                outcome, scale by 1/p: the pure collapse kernel
                                on qubit q and on ghost q + n (each scales by
                                1/sqrt(p)) (src/measurement.lisp:43-69)
    MEASURE without target      drop the coherences of q: the diagonal
                                [1, 0, 0, 1] on (q, q + n)
                                (src/measurement.lisp:111-120)
    measure-all                 qubits n-1 .. 0 measured one by one
                                (naive-measure-all, src/measurement.lisp:153-163)

Noisy gate definitions (Kraus lists) and readout POVMs follow
src/density-qvm.lisp:114-181; gate modifiers are refused there and here.
"""
from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np

from . import gates as G
from . import quil as Q
from .noise import check_kraus_ops, check_povm, perturb_measurement
from .qvm import QVMError, _UniformStream, transition_classical
from .state import HipPureState


class DensityQVM:
    """density-qvm: n qubits, vec(rho) in a 2n-qubit state (|0><0| at start)."""

    def __init__(self, num_qubits: int, dtype: str = "f64", rng=None,
                 noisy_gate_definitions: Optional[Dict] = None, readout_povms: Optional[Dict] = None):
        self.number_of_qubits = int(num_qubits)
        self.state = HipPureState(2 * self.number_of_qubits, dtype=dtype)
        self.rng = rng if rng is not None else _UniformStream()
        self.program: Optional[Q.ParsedProgram] = None
        self.memory: dict = {}
        self.pc = 0
        self.draws_used = 0
        self.noisy_gate_definitions: Dict = {}
        self.readout_povms: Dict[int, tuple] = {}
        self._pristine: Optional[HipPureState] = None
        self._sum: Optional[HipPureState] = None
        for (name, qubits), ops in (noisy_gate_definitions or {}).items():
            self.set_noisy_gate(name, qubits, ops)
        for q, povm in (readout_povms or {}).items():
            self.set_readout_povm(q, povm)

    # -- configuration ----------------------------------------------------
    def set_noisy_gate(self, gate_name: str, qubits, kraus_ops) -> None:
        ops = [np.asarray(k, dtype=np.complex128) for k in kraus_ops]
        check_kraus_ops(ops)
        self.noisy_gate_definitions[(gate_name.upper(), tuple(int(q) for q in qubits))] = ops

    def set_readout_povm(self, qubit: int, povm) -> None:
        check_povm(povm)
        self.readout_povms[int(qubit)] = tuple(float(x) for x in povm)

    # -- program ----------------------------------------------------------
    def load_program(self, program) -> "DensityQVM":
        if isinstance(program, str):
            program = Q.parse_quil(program)
        if program.qubits_needed() > self.number_of_qubits:
            raise QVMError(f"Trying to load a program with {program.qubits_needed()} qubits on a QVM of "
                           f"{self.number_of_qubits} qubits")
        self.program = program
        self.memory = {}
        for d in program.declarations.values():
            self.memory[d.name] = np.zeros(d.length, dtype=np.float64 if d.type == "REAL" else np.int64)
        self.pc = 0
        return self

    def memory_ref(self, name: str, index: int = 0):
        return self.memory[name][index]

    def run(self) -> "DensityQVM":
        if self.program is None:
            raise QVMError("no program loaded")
        code = self.program.executable_code
        self.pc = 0
        while self.pc is not None and self.pc < len(code):
            self.pc = self.transition(code[self.pc])
        return self

    def _draw(self) -> float:
        self.draws_used += 1
        return float(self.rng())

    def transition(self, ins):
        if isinstance(ins, Q.GateApplication):
            if ins.modifiers:
                raise QVMError("The density QVM doesn't support gate modifiers.")
            ops = self.noisy_gate_definitions.get((ins.name.upper(), tuple(ins.qubits)))
            if ops is None:
                U = Q.resolve_gate_matrix(ins, self.program.gate_definitions if self.program else {}, self.memory)
                if G.num_qubits_of(U) != len(ins.qubits):
                    raise QVMError(f"Gate {ins.name} takes {G.num_qubits_of(U)} qubit(s), "
                                   f"applied to {len(ins.qubits)}")
                ops = [U]
            self.apply_superoperator(ops, ins.qubits)
            return self.pc + 1
        if isinstance(ins, Q.Measure):
            if ins.target is None:
                self.measure_discard(ins.qubit)
            else:
                _, bit = self.measure(ins.qubit)
                name, idx = ins.target
                # readout noise on the stored bit (src/density-qvm.lisp:161-181)
                povm = self.readout_povms.get(ins.qubit)
                self.memory[name][idx] = perturb_measurement(bit, *povm, self._draw()) if povm else bit
            return self.pc + 1
        if isinstance(ins, Q.Reset):
            if ins.qubit is None:
                self.reset_quantum_state()
            else:
                _, bit = self.measure(ins.qubit)
                if bit == 1:
                    self.apply_superoperator([G.gate_matrix("X")], [ins.qubit])
            return self.pc + 1
        if isinstance(ins, Q.Halt):
            return None
        if isinstance(ins, (Q.Nop, Q.Label)):
            return self.pc + 1
        r = transition_classical(self, ins)
        if r is not None:
            return r
        raise QVMError(f"unsupported instruction {ins!r}")

    # -- evolution --------------------------------------------------------
    def _apply_single(self, U: np.ndarray, quil_qubits) -> None:
        n = self.number_of_qubits
        self.state.apply_matrix(np.conj(U), G.nat_tuple(*quil_qubits))
        self.state.apply_matrix(U, G.nat_tuple(*[q + n for q in quil_qubits]))

    def apply_superoperator(self, kraus_ops: List[np.ndarray], quil_qubits) -> None:
        """rho -> sum_j K_j rho K_j^H on the Quil-ordered qubits."""
        ops = [np.asarray(k, dtype=np.complex128) for k in kraus_ops]
        if not ops:
            return
        if len(ops) == 1:
            self._apply_single(ops[0], quil_qubits)
            return
        if self._pristine is None:
            self._pristine = HipPureState(self.state.num_qubits, dtype=self.state.dtype)
            self._sum = HipPureState(self.state.num_qubits, dtype=self.state.dtype)
        self._pristine.copy_from(self.state)
        for j, K in enumerate(ops):
            if j:
                self.state.copy_from(self._pristine)
            self._apply_single(K, quil_qubits)
            if j == 0:
                self._sum.copy_from(self.state)
            else:
                self._sum.accumulate(self.state)
        self.state.copy_from(self._sum)

    # -- measurement ------------------------------------------------------
    def measurement_probabilities(self) -> np.ndarray:
        return self.state.density_diagonal(self.number_of_qubits)

    def excited_probability(self, q: int) -> float:
        """Sum of the diagonal over the indices with bit q set, clamped to
        [0, 1] (a trace-1 matrix can exceed 1 by a rounding step)."""
        p = self.measurement_probabilities()
        idx = np.arange(p.size)
        return float(min(1.0, max(0.0, np.sum(p[((idx >> q) & 1) == 1]))))

    def force_measurement(self, bit: int, q: int, p1: float) -> None:
        self.state.force_measure(q, bit, p1)
        self.state.force_measure(q + self.number_of_qubits, bit, p1)

    def measure(self, q: int):
        """measure (src/measurement.lisp:87-105) on the density state."""
        if not (0 <= q < self.number_of_qubits):
            raise QVMError(f"Trying to measure qubit {q} on a QVM with only {self.number_of_qubits} qubits.")
        p1 = self.excited_probability(q)
        bit = 0 if p1 == 0.0 else (1 if self._draw() <= p1 else 0)
        self.force_measurement(bit, q, p1)
        return self, bit

    def measure_discard(self, q: int) -> None:
        self.state.apply_diagonal(np.array([1, 0, 0, 1], dtype=np.complex128), [q, q + self.number_of_qubits])

    def measure_all(self):
        bits = []
        for q in range(self.number_of_qubits - 1, -1, -1):
            _, b = self.measure(q)
            bits.insert(0, b)
        out = []
        for i, c in enumerate(bits):
            povm = self.readout_povms.get(i)
            out.append(perturb_measurement(c, *povm, self._draw()) if povm else c)
        return self, out

    # -- state ------------------------------------------------------------
    def reset_quantum_state(self) -> None:
        self.state.reset_zero()

    def amplitudes(self) -> np.ndarray:
        """vec(rho) (the state elements)."""
        return self.state.amplitudes()

    def density_matrix(self) -> np.ndarray:
        d = 1 << self.number_of_qubits
        return self.amplitudes().reshape(d, d)

    def close(self) -> None:
        for s in (self._pristine, self._sum):
            if s is not None:
                s.close()
        self._pristine = self._sum = None
        self.state.close()


def make_density_qvm(num_qubits: int, **kw) -> DensityQVM:
    """make-density-qvm (src/density-qvm.lisp:66-71)."""
    return DensityQVM(num_qubits, **kw)
