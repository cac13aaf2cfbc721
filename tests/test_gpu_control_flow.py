"""Control flow and classical memory on the device-resident machines: the
reference's jump / halt / parametric-program tests
(tests/instruction-tests.lisp), a loop whose body is a compiled gate run and
whose rotation angle is updated in classical memory (interpreted and
compiled in every mode, against the oracle on the unrolled circuit), and the
same jumps on the density machine."""
import sys

import numpy as np
import pytest

from qvm_amd import _native as N
from qvm_amd import gates as G
from qvm_amd import quil as Q

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def QV():
    from qvm_amd import qvm
    assert N.device_count() >= 1, "no HIP device visible"
    return qvm


def p1(q, k):
    a = q.amplitudes()
    idx = np.arange(a.size)
    return float(np.sum(np.abs(a[((idx >> k) & 1) == 1]) ** 2))


@pytest.mark.parametrize("compiled", [False, True])
def test_jumps_and_halt(QV, compiled):
    say(f"jumps and halt compiled={compiled}")
    cases = [
        # tests/instruction-tests.lisp:33-63
        (1, "X 0\nJUMP @SKIP\nX 0\nLABEL @SKIP", [1.0]),
        (1, "DECLARE ro BIT\nX 0\nMEASURE 0 ro[0]\nJUMP-WHEN @SKIP ro[0]\nX 0\nLABEL @SKIP", [1.0]),
        (1, "DECLARE ro BIT\nX 0\nLABEL @REDO\nMEASURE 0 ro[0]\nJUMP-UNLESS @SKIP ro[0]\nX 0\nJUMP @REDO\n"
            "LABEL @SKIP", [0.0]),
        # tests/instruction-tests.lisp:89-96
        (2, "X 0\nHALT\nX 1", [1.0, 0.0]),
    ]
    for n, text, probs in cases:
        q = QV.make_qvm(n)
        q.load_program(text)
        if compiled:
            q.compile_loaded_program(mode="tiled")
        q.run()
        for k, want in enumerate(probs):
            assert p1(q, k) == want, (text, k)
        q.close()


@pytest.mark.parametrize("compiled", [False, True])
def test_parameterized_programs(QV, compiled):
    # tests/instruction-tests.lisp:106-131
    for text in ("DECLARE ro BIT[1]\nDECLARE alpha REAL\nMOVE alpha[0] pi\nRX(alpha) 0\nMEASURE 0 ro[0]",
                 "DECLARE ro BIT[1]\nDECLARE alpha REAL\nMOVE alpha[0] (pi/2.0)\nRX(2.0*alpha) 0\nMEASURE 0 ro[0]"):
        q = QV.make_qvm(1)
        q.load_program(text)
        if compiled:
            q.compile_loaded_program(mode="plain")
        q.run()
        assert q.memory_ref("ro", 0) == 1
        q.close()


def test_measure_chain_compilation_run(QV):
    # tests/instruction-tests.lisp:186-233: chains over qubits 0..2 run on 4 qubits
    text = """DECLARE ro BIT[3]
MEASURE 0
X 0
MEASURE 0
MEASURE 0
MEASURE 2
X 0
MEASURE 0 ro[0]
MEASURE 1
MEASURE 2
X 0
MEASURE 0 ro[2]
MEASURE 0 ro[1]
MEASURE 1 ro[2]
MEASURE 2
X 0"""
    p = Q.parse_quil(text)
    code = QV.compile_measure_chains(p.executable_code, 3)
    kinds = [type(c).__name__ for c in code]
    assert kinds == ["Measure", "GateApplication", "Measure", "Measure", "Measure", "GateApplication",
                     "MeasureAll", "Nop", "Nop", "GateApplication", "MeasureAll", "Nop", "Nop", "Nop",
                     "GateApplication"]
    p.executable_code = code
    q = QV.make_qvm(4)
    q.load_program(p)
    q.run()
    assert [int(q.memory_ref("ro", k)) for k in range(3)] == [0, 1, 0]
    q.close()


LOOP = """DECLARE ro BIT
DECLARE x INTEGER
DECLARE theta REAL
MOVE theta 0.3
LABEL @top
H 0
CNOT 0 1
RZ(0.7) 2
RX(theta) 1
CPHASE(0.4) 1 3
ADD theta 0.1
ADD x 1
GT ro[0] x 3
JUMP-UNLESS @top ro[0]
H 2
SWAP 0 3"""


def _loop_reference(oracle, n):
    ref = oracle.zero_state(n)
    theta = 0.3
    for _ in range(4):
        for name, params, qs in (("H", [], [0]), ("CNOT", [], [0, 1]), ("RZ", [0.7], [2]),
                                 ("RX", [theta], [1]), ("CPHASE", [0.4], [1, 3])):
            oracle.apply_matrix(ref, G.gate_matrix(name, params), G.nat_tuple(*qs))
        theta += 0.1
    for name, qs in (("H", [2]), ("SWAP", [0, 3])):
        oracle.apply_matrix(ref, G.gate_matrix(name), G.nat_tuple(*qs))
    return ref


@pytest.mark.parametrize("mode", [None, "plain", "tiled", "fused", "auto"])
def test_loop_with_classical_updates(QV, oracle, mode):
    n = 12
    say(f"classical loop mode={mode}")
    q = QV.make_qvm(n)
    q.load_program(LOOP)
    if mode is not None:
        cp = q.compile_loaded_program(mode=mode)
        # the constant gates of the body form one compiled run; RX(theta) stays interpreted
        assert sum(isinstance(s, QV.CompiledProgram) for s in cp.steps) == 3
    q.run()
    assert q.memory_ref("x") == 4 and q.memory_ref("ro") == 1
    assert abs(q.memory_ref("theta") - 0.7) < 1e-15
    ref = _loop_reference(oracle, n)
    got = q.amplitudes()
    if mode in (None, "plain"):
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) < 1e-12
    q.close()


def test_density_machine_jumps(QV):
    from qvm_amd.density import make_density_qvm
    say("density machine jumps")
    # measure-until-zero loop: ends in |0><0| whatever the draws
    d = make_density_qvm(2)
    d.load_program("DECLARE ro BIT\nDECLARE k INTEGER\nH 0\nLABEL @REDO\nADD k 1\nMEASURE 0 ro[0]\n"
                   "JUMP-UNLESS @SKIP ro[0]\nX 0\nJUMP @REDO\nLABEL @SKIP\nX 1")
    d.run()
    rho = d.density_matrix()
    want = np.zeros((4, 4), dtype=np.complex128)
    want[2, 2] = 1
    assert np.max(np.abs(rho - want)) < 1e-12
    assert d.memory_ref("k") in (1, 2)
    d.close()


def test_program_outlives_its_state(QV):
    """A compiled program whose state is destroyed first is detached: running
    it is refused and destroying it later only frees its own graph."""
    import gc
    from qvm_amd import planner as P
    from qvm_amd.state import HipPureState
    say("program destroyed after its state")
    st = HipPureState(10, "f64")
    prog = P.build_program(st, P.plan_tiled([(G.gate_matrix("H"), [k]) for k in range(10)], 10, 8), True)
    prog.run()
    st.close()
    with pytest.raises(N.QVMHipError, match="destroyed"):
        prog.run()
    prog.close()
    # the machine closes its compiled program before the state, and a
    # dropped QVM leaves nothing that touches a freed state
    q = QV.make_qvm(10)
    q.load_program("H 0\nCNOT 0 1")
    q.compile_loaded_program(mode="tiled")
    q.run()
    q.close()
    q = QV.make_qvm(10)
    q.load_program("H 0\nCNOT 0 1")
    q.compile_loaded_program(mode="plain")
    q.state.close()
    del q
    gc.collect()
