"""Host-side control flow and classical memory: the parser's labels, jumps and
classical instructions, their transitions on classical memory, and the
segmentation of a program into compiled gate runs and interpreted steps.
Cases follow the reference's tests/instruction-tests.lisp and
tests/classical-memory-tests.lisp with the quantum part left out (the
quantum halves run on the GPU in test_gpu_control_flow.py)."""
import numpy as np
import pytest

from qvm_amd import quil as Q
from qvm_amd import qvm as QV


class ClassicalMachine:
    """pc loop over a program's classical instructions (no quantum state)."""

    def __init__(self, text):
        self.program = Q.parse_quil(text)
        self.memory = {d.name: np.zeros(d.length, dtype=np.float64 if d.type == "REAL" else np.int64)
                       for d in self.program.declarations.values()}
        self.pc = 0
        self.trace = []

    def run(self, max_steps=10000):
        code = self.program.executable_code
        self.pc = 0
        for _ in range(max_steps):
            if self.pc >= len(code):
                return self
            ins = code[self.pc]
            self.trace.append(self.pc)
            if isinstance(ins, (Q.Label, Q.Nop)):
                self.pc += 1
            elif isinstance(ins, Q.Halt):
                return self
            else:
                nxt = QV.transition_classical(self, ins)
                assert nxt is not None, ins
                self.pc = nxt
        raise AssertionError("program did not end")

    def ref(self, name, i=0):
        return self.memory[name][i]


def test_labels_patch_to_instruction_indices():
    p = Q.parse_quil("X 0\nJUMP @SKIP\nX 0\nLABEL @SKIP")
    code = p.executable_code
    assert isinstance(code[1], Q.Jump) and code[1].target == 3 and isinstance(code[3], Q.Label)
    with pytest.raises(Q.QuilParseError):
        Q.parse_quil("JUMP @nowhere")
    with pytest.raises(Q.QuilParseError):
        Q.parse_quil("DECLARE ro BIT\nJUMP-WHEN @a")


def test_move_every_type():
    # tests/instruction-tests.lisp:133-157
    m = ClassicalMachine("""DECLARE b BIT[2]
DECLARE o OCTET[2]
DECLARE int INTEGER[2]
DECLARE r REAL[2]
MOVE b[0] 1
MOVE o[0] 2
MOVE int[0] 3
MOVE r[0] 4.0
MOVE b[1] b[0]
MOVE o[1] o[0]
MOVE int[1] int[0]
MOVE r[1] r[0]""").run()
    assert [m.ref("b", i) for i in (0, 1)] == [1, 1]
    assert [m.ref("o", i) for i in (0, 1)] == [2, 2]
    assert [m.ref("int", i) for i in (0, 1)] == [3, 3]
    assert [m.ref("r", i) for i in (0, 1)] == [4.0, 4.0]
    m = ClassicalMachine("DECLARE r REAL\nDECLARE int INTEGER\nMOVE r 1.0\nMOVE int 1").run()
    assert m.ref("r") == 1 and m.ref("int") == 1
    # a constant expression as the source (tests/instruction-tests.lisp:120-131)
    m = ClassicalMachine("DECLARE alpha REAL\nMOVE alpha[0] (pi / 2.0)").run()
    assert m.ref("alpha") == np.pi / 2


@pytest.mark.parametrize("vals,expect", [
    ((0, 0, 0.0, 0.0, 0, 0, 0, 0), (0, 1, 1, 1, 0)),
    ((0, 1, 0.0, 1.0, 0, 1, 0, 1), (1, 1, 0, 0, 0)),
    ((1, 0, 1.0, 0.0, 1, 0, 1, 0), (0, 0, 0, 1, 1)),
])
def test_indirect_comparators(vals, expect):
    # tests/classical-memory-tests.lisp:145-251: LT LE EQ GE GT on each type
    text = ["DECLARE bx BIT[2]", "DECLARE rx REAL[2]", "DECLARE ix INTEGER[2]", "DECLARE ox OCTET[2]"]
    ops = ["LT", "LE", "EQ", "GE", "GT"]
    text += [f"DECLARE {o.lower()} BIT[4]" for o in ops]
    for o in ops:
        for k, src in enumerate(("bx", "rx", "ix", "ox")):
            text.append(f"{o} {o.lower()}[{k}] {src}[0] {src}[1]")
    m = ClassicalMachine("\n".join(text))
    for k, src in enumerate(("bx", "rx", "ix", "ox")):
        m.memory[src][0], m.memory[src][1] = vals[2 * k], vals[2 * k + 1]
    m.run()
    for k in range(4):
        assert tuple(int(m.ref(o.lower(), k)) for o in ops) == expect


def test_exchange():
    # tests/classical-memory-tests.lisp:253-289
    m = ClassicalMachine("""DECLARE bx BIT[2]
DECLARE rx REAL[2]
DECLARE ix INTEGER[2]
DECLARE ox OCTET[2]
MOVE bx[0] 0
MOVE bx[1] 1
MOVE rx[0] 0.0
MOVE rx[1] 55.0
MOVE ix[0] 0
MOVE ix[1] 123
MOVE ox[0] 0
MOVE ox[1] 255
EXCHANGE bx[0] bx[1]
EXCHANGE rx[0] rx[1]
EXCHANGE ix[0] ix[1]
EXCHANGE ox[0] ox[1]""").run()
    assert m.ref("bx", 1) == 0 and m.ref("rx", 1) == 0 and m.ref("ix", 1) == 0 and m.ref("ox", 1) == 0
    assert m.ref("bx", 0) == 1 and m.ref("rx", 0) == 55.0 and m.ref("ix", 0) == 123 and m.ref("ox", 0) == 255


def test_unary_arithmetic_and_logic():
    # NOT by declared type (src/transition-classical-instructions.lisp:32-51),
    # ensure-s64 wrapping, floor division, bitwise ops, CONVERT rules
    m = ClassicalMachine("""DECLARE b BIT
DECLARE o OCTET
DECLARE i INTEGER[8]
DECLARE r REAL[3]
NOT b
MOVE o 5
NOT o
MOVE i[0] 5
NOT i[0]
MOVE i[1] -7
DIV i[1] 2
MOVE i[2] 9223372036854775807
ADD i[2] 1
MOVE i[3] 12
AND i[3] 10
MOVE i[4] 12
IOR i[4] 3
MOVE i[5] 12
XOR i[5] 10
MOVE i[6] 4
NEG i[6]
MOVE r[0] 2.5
CONVERT i[7] r[0]
MOVE r[1] 7.0
DIV r[1] 2.0
MUL r[1] 3
SUB r[1] 0.5
CONVERT r[2] i[3]""").run()
    assert m.ref("b") == 1
    assert m.ref("o") == 250
    assert m.ref("i", 0) == -6
    assert m.ref("i", 1) == -4
    assert m.ref("i", 2) == -(1 << 63)
    assert m.ref("i", 3) == 8 and m.ref("i", 4) == 15 and m.ref("i", 5) == 6
    assert m.ref("i", 6) == -4
    assert m.ref("i", 7) == 2  # round half to even
    assert m.ref("r", 1) == 10.0
    assert m.ref("r", 2) == 8.0
    with pytest.raises(Q.QuilParseError):
        ClassicalMachine("DECLARE i INTEGER\nDIV i 0").run()
    with pytest.raises(Q.QuilParseError):
        ClassicalMachine("DECLARE i INTEGER\nMOVE j 1").run()
    with pytest.raises(Q.QuilParseError):
        ClassicalMachine("DECLARE i INTEGER\nMOVE i[3] 1").run()


def test_counting_loop():
    # LABEL / ADD / GT / JUMP-UNLESS: the body runs until x > 2
    m = ClassicalMachine("""DECLARE ro BIT
DECLARE x INTEGER
LABEL @top
ADD x 1
GT ro[0] x 2
JUMP-UNLESS @top ro[0]
JUMP @end
MOVE x 100
LABEL @end""").run()
    assert m.ref("x") == 3 and m.ref("ro") == 1
    assert m.trace.count(1) == 3  # ADD executed three times
    # JUMP-WHEN takes the branch on a 1 bit only
    m = ClassicalMachine("""DECLARE b BIT
DECLARE x INTEGER
JUMP-WHEN @skip b
MOVE x 1
LABEL @skip
MOVE b 1
JUMP-WHEN @out b
MOVE x 2
LABEL @out""").run()
    assert m.ref("x") == 1


def test_segmentation_renumbers_jumps():
    p = Q.parse_quil("""DECLARE ro BIT
DECLARE theta REAL
X 0
H 1
LABEL @redo
MEASURE 0 ro[0]
NOP
JUMP-UNLESS @skip ro[0]
X 0
CNOT 0 1
RX(theta) 0
JUMP @redo
LABEL @skip
H 1""")
    memory = {"ro": np.zeros(1, dtype=np.int64), "theta": np.zeros(1)}

    def gate(ins):
        if not isinstance(ins, Q.GateApplication) or QV._reads_memory(ins, memory):
            return None
        return (ins.name, list(ins.qubits))

    steps = QV.segment_program(p.executable_code, gate, lambda run: ("RUN", [g for g, _ in run]))
    kinds = [s[0] if isinstance(s, tuple) else type(s).__name__ for s in steps]
    assert kinds == ["RUN", "Label", "Measure", "Jump", "RUN", "GateApplication", "Jump", "Label", "RUN"]
    assert steps[0][1] == ["X", "H"] and steps[4][1] == ["X", "CNOT"]
    # the memory-dependent rotation stays interpreted
    assert steps[5].name == "RX"
    # jump targets point at the label steps
    assert steps[3].target == 7 and isinstance(steps[7], Q.Label) and steps[7].name == "skip"
    assert steps[6].target == 1 and isinstance(steps[1], Q.Label) and steps[1].name == "redo"
    # the parsed program is left as it was
    assert p.executable_code[5].target == 10
