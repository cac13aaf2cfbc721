"""qvm's pure-state machine API over the HBM-resident engine.

Mirrors the reference's Lisp entry points so a user of qvm finds the same
operations (names in snake_case):

    make_qvm            src/qvm.lisp:150-164
    load_program        src/classical-memory-mixin.lisp:114-159
    run                 src/execution.lisp:31-37
    run_program         src/execution.lisp:46-59
    transition          src/transition.lisp:46-194 (gate, MEASURE, RESET, HALT)
    apply_gate_to_state src/apply-gate.lisp:106-160
    measure             src/measurement.lisp:87-105
    measure_all         src/measurement.lisp:122-143
    amplitudes          src/qvm.lisp:63-82
    compile_loaded_program  src/qvm.lisp:166-175 (plain / tiled / fused / auto
                        plans; measure chains -> measure-all,
                        src/compile-gate.lisp:538-599)
    perform_expectation app/src/api/expectation.lisp:78-89 (pure state)
    wavefunction / probabilities octets  app/src/handle-request.lisp:132-175

Random draws come from ``qvm.rng`` (a callable returning uniforms in [0,1));
the reference draws from mt19937 (src/utilities.lisp:306-335), which is not
vendored, so callers that need a reproducible stream pass their own.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import gates as G
from . import quil as Q
from .state import HipPureState, Program


class QVMError(RuntimeError):
    pass


@dataclass
class MeasureAll:
    """measure-all pseudo-instruction (src/compile-gate.lisp:540-545):
    storage lists (qubit, (name, index)) for the MEASUREs of the chain that
    store their bit."""
    storage: List[Tuple[int, Tuple[str, int]]]


def _measure_chain_at(i: int, code) -> int:
    """Length of the run of measurements starting at i
    (measure-chain-at, src/compile-gate.lisp:554-563)."""
    n = 0
    for ins in code[i:]:
        if not isinstance(ins, Q.Measure):
            break
        n += 1
    return n


def compile_measure_chains(code, n: int) -> list:
    """Replace every run of MEASUREs that covers all qubits 0..n-1 by one
    MeasureAll followed by NOPs for the rest of the run
    (find-measure-chains / compile-measure-chains,
    src/compile-gate.lisp:565-599)."""
    code = list(code)
    i = 0
    chains = []
    while i < len(code):
        c = _measure_chain_at(i, code)
        if c > 0:
            if c >= n:
                bits = 0
                for m in code[i:i + c]:
                    bits |= 1 << m.qubit
                if bits == (1 << n) - 1:
                    chains.append((i, code[i:i + c]))
            i += c
        else:
            i += 1
    for start, chain in chains:
        for k in range(start, start + len(chain)):
            code[k] = Q.Nop()
        code[start] = MeasureAll([(m.qubit, m.target) for m in chain if m.target is not None])
    return code


def transition_classical(qvm, ins) -> Optional[int]:
    """Jumps and classical instructions on a machine with ``pc``, ``memory``
    and ``program`` (src/transition.lisp:98-115: JUMP sets the pc to the
    label, JUMP-WHEN when the bit is 1, JUMP-UNLESS when it is 0;
    src/transition-classical-instructions.lisp).  Returns the next pc, or
    None when ``ins`` is neither."""
    if isinstance(ins, Q.Jump):
        if ins.when is None:
            return ins.target
        name, idx = ins.cond
        if name not in qvm.memory:
            raise QVMError(f"undeclared memory {name}")
        taken = int(qvm.memory[name][idx]) == 1 if ins.when else int(qvm.memory[name][idx]) == 0
        return ins.target if taken else qvm.pc + 1
    if isinstance(ins, Q.Classical):
        types = {d.name: d.type for d in qvm.program.declarations.values()} if qvm.program else None
        Q.execute_classical(ins, qvm.memory, types)
        return qvm.pc + 1
    return None


def segment_program(code, gate, compile_run) -> list:
    """Cut an instruction vector into steps: each maximal run of
    instructions for which ``gate(ins)`` returns a (matrix, qubits) pair
    becomes one ``compile_run(list_of_pairs)`` step; NOPs are dropped; every
    other instruction is its own step.  Labels are steps of their own, so a
    run never spans one and a jump always lands on a step boundary; jump
    targets (instruction indices of labels) are renumbered to step indices."""
    steps: list = []
    run: list = []
    label_step = {}
    jumps = []

    def flush():
        if run:
            steps.append(compile_run(list(run)))
            run.clear()

    for i, ins in enumerate(code):
        g = gate(ins)
        if g is not None:
            run.append(g)
        elif isinstance(ins, Q.Nop):
            continue
        else:
            flush()
            if isinstance(ins, Q.Label):
                label_step[i] = len(steps)
            elif isinstance(ins, Q.Jump):
                jumps.append(len(steps))
            steps.append(ins)
    flush()
    for j in jumps:
        J = steps[j]
        if J.target not in label_step:
            raise QVMError(f"jump to @{J.label} does not land on a label")
        steps[j] = Q.Jump(J.label, J.when, J.cond, label_step[J.target])
    return steps


class _MemoryProbe(dict):
    """Stands in for classical memory while a gate's parameters are
    evaluated, recording whether any of them reads it."""

    def __init__(self, memory: dict):
        super().__init__()
        self.memory = memory
        self.read = False

    def __getitem__(self, name):
        self.read = True
        return self.memory[name]


def _reads_memory(ins: "Q.GateApplication", memory: dict) -> bool:
    probe = _MemoryProbe(memory)
    for p in ins.params:
        p({}, probe)
    return probe.read


class _UniformStream:
    def __init__(self, seed=None):
        self._g = np.random.default_rng(seed)

    def __call__(self) -> float:
        return float(self._g.random())


class PureStateQVM:
    """pure-state-qvm (src/qvm.lisp:106-148) with an HBM-resident state."""

    def __init__(self, num_qubits: int, dtype="f64", rng: Optional[Callable[[], float]] = None,
                 device: int = -1, memory_cap_bytes: int = 0, compiled_measurement: bool = False):
        self.number_of_qubits = int(num_qubits)
        self.state = HipPureState(num_qubits, dtype=dtype, device=device, memory_cap_bytes=memory_cap_bytes)
        self.rng = rng if rng is not None else _UniformStream()
        self.program: Optional[Q.ParsedProgram] = None
        self.memory: dict = {}
        self.pc = 0
        self.compiled_measurement = compiled_measurement
        self.draws_used = 0
        self._compiled: Optional["CompiledProgram"] = None

    # -- program ---------------------------------------------------------
    def load_program(self, program) -> "PureStateQVM":
        """load-program: qubit-count check, classical memory, pc reset."""
        if isinstance(program, str):
            program = Q.parse_quil(program)
        need = program.qubits_needed()
        if need > self.number_of_qubits:
            raise QVMError(f"Trying to load a program with {need} qubits on a QVM of "
                           f"{self.number_of_qubits} qubits")
        self.program = program
        self.memory = {}
        for d in program.declarations.values():
            dt = np.float64 if d.type == "REAL" else np.int64
            self.memory[d.name] = np.zeros(d.length, dtype=dt)
        self.pc = 0
        self._drop_compiled()
        return self

    def _drop_compiled(self) -> None:
        if self._compiled is not None:
            self._compiled.close()
            self._compiled = None

    def memory_ref(self, name: str, index: int = 0):
        return self.memory[name][index]

    def reset_classical_memory(self) -> None:
        for v in self.memory.values():
            v[:] = 0

    def reset_quantum_state(self) -> None:
        """reset-quantum-state (src/qvm.lisp:84-87)."""
        self.state.reset_zero()

    # -- execution -------------------------------------------------------
    def run(self) -> "PureStateQVM":
        if self.program is None:
            raise QVMError("no program loaded")
        if self._compiled is not None:
            self._compiled.run(self)
            self.pc = len(self.program.executable_code)
            return self
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
            U = Q.resolve_gate_matrix(ins, self.program.gate_definitions if self.program else {}, self.memory)
            k = G.num_qubits_of(U)
            if k != len(ins.qubits):
                raise QVMError(f"Gate {ins.name} takes {k} qubit(s), applied to {len(ins.qubits)}")
            apply_gate_to_state(self.state, U, ins.qubits)
            return self.pc + 1
        if isinstance(ins, Q.Measure):
            _, bit = self.measure(ins.qubit)
            if ins.target is not None:
                name, idx = ins.target
                self.memory[name][idx] = bit
            return self.pc + 1
        if isinstance(ins, Q.Reset):
            if ins.qubit is None:
                self.reset_quantum_state()
            else:
                # src/transition.lisp:76-95: measure, then X if the result was 1
                _, bit = self.measure(ins.qubit)
                if bit == 1:
                    apply_gate_to_state(self.state, G.gate_matrix("X"), [ins.qubit])
            return self.pc + 1
        if isinstance(ins, MeasureAll):
            # transition on measure-all (src/transition.lisp:130-135)
            _, bits = self.measure_all()
            for q, (name, idx) in ins.storage:
                self.memory[name][idx] = bits[q]
            return self.pc + 1
        if isinstance(ins, Q.Halt):
            return None
        if isinstance(ins, (Q.Nop, Q.Label)):
            return self.pc + 1
        r = transition_classical(self, ins)
        if r is not None:
            return r
        raise QVMError(f"unsupported instruction {ins!r}")

    # -- measurement -----------------------------------------------------
    def measure(self, q: int):
        """measure: interpreted rule (no draw when p1 == 0) unless
        ``compiled_measurement`` selects the compiled projector rule."""
        if not (0 <= q < self.number_of_qubits):
            raise QVMError(f"Trying to measure qubit {q} on a QVM with only {self.number_of_qubits} qubits.")
        if self.compiled_measurement:
            u = self._draw()
            bit, _ = self.state.measure(q, u, compiled=True)
            return self, bit
        p1 = self.state.prob_excited(q)
        if p1 == 0.0:
            bit = 0
        else:
            bit = 1 if self._draw() <= p1 else 0
        self.state.force_measure(q, bit, p1)
        return self, bit

    def measure_all(self):
        """measure-all: one draw, collapse to the sampled basis state, bits
        LSB first."""
        b = self.state.measure_all(self._draw())
        return self, [(b >> i) & 1 for i in range(self.number_of_qubits)]

    # -- amplitudes ------------------------------------------------------
    def amplitudes(self) -> np.ndarray:
        return self.state.amplitudes()

    def set_amplitudes(self, amps) -> None:
        self.state.write_amplitudes(amps)

    def nth_amplitude(self, i: int) -> complex:
        return complex(self.state.read_amplitudes(i, 1)[0])

    # -- compilation -----------------------------------------------------
    def compile_loaded_program(self, mode: str = "tiled", use_graph: bool = True,
                               measure_chains: bool = True, **kw) -> "CompiledProgram":
        """compile-loaded-program (src/qvm.lisp:166-175).  Measure chains over
        all qubits become one measure-all (*compile-measure-chains*); every
        maximal run of gate applications with constant parameters is planned
        into engine launches (plain: one launch per gate; tiled: LDS tile
        passes; fused: dense MFMA blocks; auto: the plan the trained
        launch-cost model prices lowest) and captured.  Gates whose
        parameters read classical memory stay interpreted, as
        compile-instruction leaves them (src/compile-gate.lisp:484-489), and
        so do the other instructions.  Labels end a run, so a jump always
        lands on a step boundary; jump targets are renumbered to steps."""
        from . import planner
        if self.program is None:
            raise QVMError("no program loaded")
        code = self.program.executable_code
        if measure_chains:
            code = compile_measure_chains(code, self.number_of_qubits)

        def gate(ins):
            if not isinstance(ins, Q.GateApplication) or _reads_memory(ins, self.memory):
                return None
            U = Q.resolve_gate_matrix(ins, self.program.gate_definitions, self.memory)
            if G.num_qubits_of(U) != len(ins.qubits):
                raise QVMError(f"Gate {ins.name} takes {G.num_qubits_of(U)} qubit(s), "
                               f"applied to {len(ins.qubits)}")
            return (U, list(ins.qubits))

        steps = segment_program(code, gate, lambda run: planner.compile_gates(
            self.state, run, mode=mode, use_graph=use_graph, **kw))
        self._drop_compiled()
        self._compiled = CompiledProgram.from_steps(steps, mode)
        return self._compiled

    # -- wire formats (app/src/handle-request.lisp:132-175) ---------------
    def wavefunction_octets(self) -> bytes:
        """The wavefunction reply: every amplitude as big-endian re, im
        doubles (16 octets each, octets-required-for-quantum-state)."""
        return self.state.wavefunction_octets()

    def probabilities_octets(self) -> bytes:
        """The probabilities reply: |psi_i|^2 as big-endian doubles."""
        return self.state.probabilities_octets()

    def close(self) -> None:
        """Release the compiled program's graphs and operands, then the state."""
        self._drop_compiled()
        self.state.close()


def make_qvm(num_qubits: int, **kw) -> PureStateQVM:
    return PureStateQVM(num_qubits, **kw)


def run_program(num_qubits: int, program, **kw) -> PureStateQVM:
    q = make_qvm(num_qubits, **kw)
    q.load_program(program)
    return q.run()


def apply_gate_to_state(state: HipPureState, U: np.ndarray, quil_qubits) -> None:
    """Pure-state apply-gate-to-state: matrix in Quil convention, qubits in
    Quil argument order; the engine receives the nat-tuple
    (src/apply-gate.lisp:109-160)."""
    state.apply_matrix(U, G.nat_tuple(*quil_qubits))


class CompiledProgram:
    """Handle returned by compile_loaded_program: captured gate programs and
    the non-gate instructions between them, in program order."""

    def __init__(self, program: Optional[Program], plan: list, mode: str):
        self.program = program
        self.plan = plan
        self.mode = mode
        self.steps: list = [self] if program is not None else []

    @classmethod
    def from_steps(cls, steps: list, mode: str) -> "CompiledProgram":
        cp = cls(None, [], mode)
        cp.steps = steps
        cp.plan = [it for st in steps if isinstance(st, CompiledProgram) for it in st.plan]
        return cp

    def run(self, qvm: Optional["PureStateQVM"] = None) -> None:
        if self.program is not None:
            self.program.run()
            return
        pc = 0
        while pc < len(self.steps):
            st = self.steps[pc]
            if isinstance(st, CompiledProgram):
                st.run()
                pc += 1
                continue
            if qvm is None:
                raise QVMError("this compiled program has non-gate instructions; run it through its QVM")
            qvm.pc = pc  # transition returns the next step: pc + 1, or a jump's step
            nxt = qvm.transition(st)
            if nxt is None:
                return  # HALT
            pc = nxt

    def close(self) -> None:
        if self.program is not None:
            self.program.close()
        for st in self.steps:
            if isinstance(st, CompiledProgram) and st is not self:
                st.close()


def perform_expectation(state_prep, operators, num_qubits: int, dtype: str = "f64", **kw) -> List[float]:
    """Expectation values <prep|O|prep> of operator programs on one prepared
    state (pure-state branch of perform-expectation,
    app/src/api/expectation.lisp:42-89): run the state preparation, keep a
    device copy of the prepared state, and for each operator program restore
    the copy, run the program and take the inner product with the prepared
    state (a device reduction).  Returns the real parts; the reference
    asserts |imag| < 1e-14."""
    q = make_qvm(num_qubits, dtype=dtype, **kw)
    q.load_program(state_prep)
    q.run()
    prepared = HipPureState(num_qubits, dtype=dtype)
    prepared.copy_from(q.state)
    out = []
    tol = 1e-14 if dtype == "f64" else 1e-6
    try:
        for i, op in enumerate(operators):
            if i > 0:
                q.state.copy_from(prepared)
            q.load_program(op)
            q.run()
            v = prepared.inner_product(q.state)
            if abs(v.imag) >= tol:
                raise QVMError(f"expectation value with imaginary part {v.imag!r}")
            out.append(float(v.real))
    finally:
        prepared.close()
        q.close()
    return out
