"""CPU tests of the host side: Quil front end, planners, and the C-ABI library
surface (loads, exports every declared symbol, rejects bad arguments before
touching a device).  No kernel is launched here."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from qvm_amd import _native as N
from qvm_amd import circuits as CI
from qvm_amd import gates as G
from qvm_amd import planner as P
from qvm_amd import quil as Q

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def full_unitary(U, quil_qubits, n):
    from tests.test_oracle import kron_unitary
    return kron_unitary(U, quil_qubits, n)


def simulate(gates, n, psi):
    for U, q in gates:
        psi = full_unitary(U, q, n) @ psi
    return psi


def emulate_tile_op(op, tq, n, psi):
    """Numpy model of the tile kernel's op semantics on the full state."""
    g = lambda local: tq[local]  # noqa: E731
    m = np.array([op.m[i] for i in range(32)]).view(np.complex128)
    ctrl = [g(b) for b in range(32) if (op.ctrl_mask >> b) & 1]
    if op.kind == N.TILE_DENSE1:
        U, qs = m[:4].reshape(2, 2), [g(op.q0)]
    elif op.kind == N.TILE_DIAG1:
        U, qs = np.diag(m[:2]), [g(op.q0)]
    elif op.kind == N.TILE_DENSE2:
        U, qs = m[:16].reshape(4, 4), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_DIAG2:
        U, qs = np.diag(m[:4]), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_CPERM:
        U, qs = G.gate_matrix("CNOT"), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_SWAP:
        U, qs = G.gate_matrix("SWAP"), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_CPHASE:
        U, qs = np.diag([1, 1, 1, m[0]]), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_GDIAG:
        # diagonal over (tile qubit q0, outside qubits g0[, g1]): entry
        # index b(q0) + 2 b(g0) + 4 b(g1); Quil order is the reverse
        ng = op.q1 >> 16
        outside = [op.q1 & 0xFF, (op.q1 >> 8) & 0xFF][:ng]
        U, qs = np.diag(m[:1 << (1 + ng)]), list(reversed([g(op.q0)] + outside))
    else:
        raise AssertionError(op.kind)
    for c in ctrl:
        U, qs = G.controlled(U), [c] + qs
    return full_unitary(U, qs, n) @ psi


def emulate_plan(plan, n, psi):
    for it in plan:
        if it.kind in ("gate", "fused"):
            # engine order -> Quil order
            psi = full_unitary(it.U, list(reversed(it.qubits)), n) @ psi
        else:
            for op in it.ops:
                psi = emulate_tile_op(op, it.qubits, n, psi)
    return psi


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


# --- Quil front end ----------------------------------------------------------

def test_parse_program_and_expressions():
    p = Q.parse_quil("""
DECLARE ro BIT[2]
DECLARE theta REAL[2]
# comment
H 0; CNOT 0 1
RX(pi/2) 1
RZ(-2*theta[1] + 0.5) 0
CPHASE(pi/2^3) 0 1
DAGGER CONTROLLED RY(2*pi/3) 0 1
MEASURE 0 ro[1]
RESET 1
RESET
HALT
""")
    kinds = [type(i).__name__ for i in p.executable_code]
    assert kinds == ["GateApplication"] * 6 + ["Measure", "Reset", "Reset", "Halt"]
    assert p.declarations["theta"].length == 2
    mem = {"theta": np.array([0.0, 0.25])}
    rz = p.executable_code[3]
    assert abs(rz.params[0]({}, mem) - 0.0) < 1e-15
    cp = p.executable_code[4]
    assert abs(cp.params[0]({}, mem) - math.pi / 8) < 1e-15
    U = Q.resolve_gate_matrix(p.executable_code[5], {}, mem)
    want = G.dagger(G.controlled(G.gate_matrix("RY", [2 * math.pi / 3])))
    assert np.allclose(U, want)
    assert p.qubits_needed() == 2


def test_defgate_matrix_and_permutation():
    p = Q.parse_quil("""DEFGATE G(%a):
    cos(%a), sin(%a)
    -sin(%a), cos(%a)
DEFGATE CYC AS PERMUTATION:
    1, 2, 3, 0
G(0.0) 0
CYC 0 1
""")
    U = Q.resolve_gate_matrix(p.executable_code[0], p.gate_definitions, {})
    assert np.allclose(U, np.eye(2))
    V = Q.resolve_gate_matrix(p.executable_code[1], p.gate_definitions, {})
    assert np.array_equal(V @ np.eye(4)[:, 0], np.eye(4)[:, 1])  # |0> -> |1>


def test_parse_errors():
    with pytest.raises(Q.QuilParseError):
        Q.parse_quil("DECLARE x FOO[2]")
    with pytest.raises(Q.QuilParseError):
        Q.resolve_gate_matrix(Q.parse_quil("NOTAGATE 0").executable_code[0], {}, {})


def test_changed_amplitudes_definition():
    n = 10
    assert G.changed_amplitudes(G.gate_matrix("H"), n) == 1 << n
    assert G.changed_amplitudes(G.gate_matrix("RZ", [0.3]), n) == 1 << n
    assert G.changed_amplitudes(G.gate_matrix("CNOT"), n) == 1 << (n - 1)
    assert G.changed_amplitudes(G.gate_matrix("SWAP"), n) == 1 << (n - 1)
    assert G.changed_amplitudes(G.gate_matrix("CZ"), n) == 1 << (n - 2)
    assert G.changed_amplitudes(G.gate_matrix("CPHASE", [0.1]), n) == 1 << (n - 2)
    g = CI.layered(20, 20, 1)
    assert CI.amp_updates(g, 20) == 938475520  # 800 1q + 190 CNOT at n=20


def test_workload_shapes():
    assert len(CI.layered(30, 40, 1)) == 2400 + 580
    q = CI.qft(30)
    names = [G.num_qubits_of(U) for U, _ in q]
    assert names.count(1) == 30 and names.count(2) == 435 + 15
    h = CI.haar_cz(30, 20, 3)
    assert len(h) == 600 + 300
    for U, _ in h[:30]:
        assert G.is_unitary(U, 1e-12)
    inst = CI.qaoa_instances(8, seed=4)
    for n, edges, gam, bet in inst:
        deg = np.zeros(n, int)
        for a, b in edges:
            deg[a] += 1
            deg[b] += 1
        assert (deg == 3).all() and 1 <= len(gam) <= 4 and len(gam) == len(bet)
    u = CI.uniforms(7, 1000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.05


# --- planners ----------------------------------------------------------------

@pytest.mark.parametrize("merge", [False, True])
def test_tiled_plan_is_the_same_unitary(merge):
    n = 9
    gates = CI.layered(n, 5, seed=2) + CI.qft(n)[:20] + [(G.gate_matrix("CCNOT"), [0, 4, 8])]
    gates += [(G.gate_matrix("RZZ", [0.4]), [2, 7]), (G.gate_matrix("SWAP"), [8, 1]),
              (G.controlled(G.gate_matrix("RY", [0.3])), [6, 2])]
    plan = P.plan_tiled(gates, n, m=8, low=3, merge_1q=merge)
    assert any(it.kind == "tile" for it in plan)
    psi0 = rand_state(n, 1)
    want = simulate(gates, n, psi0)
    got = emulate_plan(plan, n, psi0)
    assert np.max(np.abs(got - want)) < 1e-12
    covered = sorted(i for it in plan for i in it.gates)
    if not merge:
        assert covered == list(range(len(gates)))


@pytest.mark.parametrize("kmax", [2, 3, 4, 5])
def test_fused_plan_is_the_same_unitary(kmax):
    n = 8
    gates = CI.random_mix(n, 6, seed=kmax)
    plan = P.plan_fused(gates, n, kmax)
    assert all(len(it.qubits) <= kmax for it in plan)
    psi0 = rand_state(n, 2)
    assert np.max(np.abs(emulate_plan(plan, n, psi0) - simulate(gates, n, psi0))) < 1e-12
    assert len(plan) < len(gates)


def test_embed_matches_kron():
    rng = np.random.default_rng(3)
    U = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    block = [5, 1, 3]  # index bit j <-> block[j]
    E = P.embed(U, [3, 5], block)
    # compare on a 6-qubit state via full unitaries
    n = 6
    full = full_unitary(E, list(reversed(block)), n)
    assert np.allclose(full, full_unitary(U, [3, 5], n))


# --- C ABI surface -----------------------------------------------------------

def declared_symbols():
    text = open(os.path.join(REPO, "include", "qvmhip.h")).read()
    return sorted(set(re.findall(r"QVMHIP_API\s+[\w\s\*]+?\b(qvmhip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 35
    L = N.lib()
    for s in syms:
        assert hasattr(L, s), s
        assert s in N.SIGNATURES, f"{s} has no ctypes signature"
    assert N.lib().qvmhip_version().decode().startswith("qvmhip")


def test_error_paths_without_a_device():
    L = N.lib()
    h = C.c_void_p()
    cfg = N.Config(C.sizeof(N.Config), -1, 1 << 20, 0)  # 1 MiB cap
    assert L.qvmhip_create(24, N.F64, C.byref(cfg), C.byref(h)) == N.ENOMEM
    assert "cap" in N.last_error()
    assert L.qvmhip_create(0, N.F64, None, C.byref(h)) == N.EINVAL
    assert L.qvmhip_create(10, 7, None, C.byref(h)) == N.EINVAL
    bad = N.Config(3, -1, 0, 0)
    assert L.qvmhip_create(10, N.F64, C.byref(bad), C.byref(h)) == N.EINVAL
    nbytes = C.c_uint64()
    assert L.qvmhip_state_bytes(30, N.F64, C.byref(nbytes)) == N.OK
    assert nbytes.value >= 16 << 30
    assert L.qvmhip_apply_matrix(None, None, 1, None) == N.EINVAL
    assert L.qvmhip_measure(None, 0, 0.5, 0, None, None) == N.EINVAL
    assert L.qvmhip_read_wavefunction_be(None, 0, 1, None) == N.EINVAL
    op = N.TileOp()
    op.kind, op.q0 = N.TILE_DENSE1, 0
    for dtype, m in (("f64", 5), ("f64", 13), ("f32", 14)):
        with pytest.raises(N.QVMHipError):
            N.tile_schedule(dtype, m, [op])
    op.q0 = 12
    with pytest.raises(N.QVMHipError):
        N.tile_schedule("f64", 12, [op])  # local bit outside the tile
    assert L.qvmhip_batch_set_flags(None, 0) == N.EINVAL
    assert L.qvmhip_total_probability(None, None) == N.EINVAL
    # GDIAG: one or two outside qubits, no controls, local q0 inside the tile
    g = N.TileOp()
    g.kind, g.q0, g.q1 = N.TILE_GDIAG, 3, 20 | (1 << 16)
    assert len(N.tile_schedule("f64", 11, [g])) == 1
    for q0, q1, ctrl in ((3, 20, 0), (3, 20 | (3 << 16), 0), (11, 20 | (1 << 16), 0), (3, 20 | (1 << 16), 1)):
        g.q0, g.q1, g.ctrl_mask = q0, q1, ctrl
        with pytest.raises(N.QVMHipError):
            N.tile_schedule("f64", 11, [g])
    g.kind = 20  # beyond the public kinds
    with pytest.raises(N.QVMHipError):
        N.tile_schedule("f64", 11, [g])


# --- compiled measure chains and noise helpers (host logic) ------------------

def test_measure_chains_compile_to_measure_all():
    from qvm_amd import qvm as QV
    p = Q.parse_quil("DECLARE ro BIT[3]\nH 0\nMEASURE 2 ro[2]\nMEASURE 0\nMEASURE 1 ro[0]\nX 1\n"
                     "MEASURE 0 ro[1]\nMEASURE 1 ro[2]")
    code = QV.compile_measure_chains(p.executable_code, 3)
    # the first run covers qubits {0, 1, 2}: one MeasureAll then NOPs; the
    # second run covers only {0, 1} and stays as it is
    assert isinstance(code[1], QV.MeasureAll)
    assert code[1].storage == [(2, ("ro", 2)), (1, ("ro", 0))]  # MEASURE 0 (discard) stores nothing
    assert isinstance(code[2], Q.Nop) and isinstance(code[3], Q.Nop)
    assert isinstance(code[5], Q.Measure) and isinstance(code[6], Q.Measure)
    # a run longer than n that still covers every qubit is one chain
    p2 = Q.parse_quil("MEASURE 0\nMEASURE 1\nMEASURE 0")
    c2 = QV.compile_measure_chains(p2.executable_code, 2)
    assert isinstance(c2[0], QV.MeasureAll) and all(isinstance(x, Q.Nop) for x in c2[1:])
    # too short: untouched
    assert QV.compile_measure_chains(p2.executable_code[:1], 2)[0] == p2.executable_code[0]


def test_noise_helpers():
    from qvm_amd import noise as NZ
    ops = NZ.make_pauli_noise_map(0.1, 0.0, 0.25)
    assert len(ops) == 4
    NZ.check_kraus_ops(ops)
    np.testing.assert_allclose(ops[0], np.sqrt(0.65) * np.eye(2))
    np.testing.assert_allclose(ops[3], np.sqrt(0.25) * G.gate_matrix("Z"))
    assert len(NZ.make_pauli_noise_map(0.5, 0.25, 0.25)) == 3  # no identity term at sum 1
    g = NZ.make_pauli_perturbed_1q_gate("H", 0.1, 0.1, 0.1)
    NZ.check_kraus_ops(g)
    np.testing.assert_allclose(g[1], np.sqrt(0.1) * G.gate_matrix("X") @ G.gate_matrix("H"))
    with pytest.raises(Exception):
        NZ.check_kraus_ops([np.eye(2), np.eye(2)])
    with pytest.raises(Exception):
        NZ.make_pauli_noise_map(0.6, 0.6, 0.0)
    NZ.check_povm((0.9, 0.2, 0.1, 0.8))
    with pytest.raises(Exception):
        NZ.check_povm((0.9, 0.2, 0.2, 0.8))
    # perturb-measurement (src/channel-qvm.lisp:185-195): 0 iff r <= p(0|actual)
    assert NZ.perturb_measurement(0, 0.9, 0.2, 0.1, 0.8, 0.9) == 0
    assert NZ.perturb_measurement(0, 0.9, 0.2, 0.1, 0.8, 0.91) == 1
    assert NZ.perturb_measurement(1, 0.9, 0.2, 0.1, 0.8, 0.2) == 0
    assert NZ.perturb_measurement(1, 0.9, 0.2, 0.1, 0.8, 0.21) == 1


# --- tile scheduler: the slot-space program equals the op list ---------------

def _apply_vec(X, op, m, tb=None, f32=False, gbits=None):
    """Apply one tile op (public kinds in local bits, or the scheduler's
    slot-space kinds when tb is given) to a 2^m vector.  f32: the scheduled
    FP32 form, whose operands are stored as floats.  gbits: the values of
    the qubits outside the tile (GDIAG), state qubit -> bit."""
    S = 1 << m
    idx = np.arange(S)
    mm = np.array([op.m[i] for i in range(32)])
    if f32 and op.kind not in (N.TILE_REMAP, N.TILE_RSWAP, N.TILE_XREG):
        mm = np.frombuffer(mm.tobytes(), dtype=np.float32)[:32].astype(np.float64)
    cm = op.ctrl_mask
    ok = (idx & cm) == cm
    k = op.kind
    if k == N.TILE_REMAP:
        src = np.frombuffer(mm.tobytes()[:16], dtype=np.uint8)
        old = np.zeros(S, dtype=np.int64)
        for j in range(m):
            old |= ((idx >> j) & 1) << int(src[j])
        return X[old]
    if k == N.TILE_RDIAG:  # slot s is in register s >> tb
        f = mm[:32].view(np.complex128) if not f32 else mm[:32].reshape(-1, 2) @ np.array([1, 1j])
        return X * f[idx >> tb]
    if k == N.TILE_RSWAP:
        A, B = tb + op.q0, tb + op.q1
        sw = idx.copy()
        diff = ((idx >> A) & 1) != ((idx >> B) & 1)
        sw[diff] ^= (1 << A) | (1 << B)
        return X[sw]
    if k in (N.TILE_D1REAL, N.TILE_D1RX):  # reduced uncontrolled 1q on a register bit
        assert op.q0 >= tb and cm == 0
        a_, b_, c_, d_ = mm[:4]
        u = np.array([[a_, b_], [c_, d_]], dtype=np.complex128) if k == N.TILE_D1REAL else \
            np.array([[a_, 1j * b_], [1j * c_, d_]])
        b = 1 << op.q0
        a = idx[(idx & b) == 0]
        X = X.copy()
        xa, xb = X[a].copy(), X[a | b].copy()
        X[a], X[a | b] = u[0, 0] * xa + u[0, 1] * xb, u[1, 0] * xa + u[1, 1] * xb
        return X
    X = X.copy()
    if k == N.TILE_DENSE1:
        if tb is not None:
            assert op.q0 >= tb
        u = mm[:8].view(np.complex128).reshape(2, 2)
        b = 1 << op.q0
        a = idx[((idx & b) == 0) & ok]
        xa, xb = X[a].copy(), X[a | b].copy()
        X[a], X[a | b] = u[0, 0] * xa + u[0, 1] * xb, u[1, 0] * xa + u[1, 1] * xb
    elif k == N.TILE_DENSE2:
        if tb is not None:
            assert op.q0 == tb and op.q1 == tb + 1
        u = mm.view(np.complex128).reshape(4, 4)
        b0, b1 = 1 << op.q0, 1 << op.q1
        a = idx[((idx & (b0 | b1)) == 0) & ok]
        y = u @ np.stack([X[a], X[a | b0], X[a | b1], X[a | b0 | b1]])
        X[a], X[a | b0], X[a | b1], X[a | b0 | b1] = y
    elif k in (N.TILE_XREG, N.TILE_CPERM):
        if tb is not None:
            assert k == N.TILE_XREG and op.q0 >= tb
        b = 1 << op.q0
        sel = ((idx & b) == 0) & ok
        if k == N.TILE_CPERM:
            sel &= ((idx >> op.q1) & 1) == 1
        a = idx[sel]
        X[a], X[a | b] = X[a | b].copy(), X[a].copy()
    elif k == N.TILE_SWAP:
        if tb is not None:
            assert {op.q0, op.q1} == {tb, tb + 1}
        b0, b1 = 1 << op.q0, 1 << op.q1
        a = idx[((idx & (b0 | b1)) == 0) & ok]
        X[a | b0], X[a | b1] = X[a | b1].copy(), X[a | b0].copy()
    elif k in (N.TILE_DIAG1, N.TILE_DIAG2):
        d = mm[:8].view(np.complex128)
        j = (idx >> op.q0) & 1
        if k == N.TILE_DIAG2:
            j = j | (((idx >> op.q1) & 1) << 1)
        X = np.where(ok, X * d[j], X)
    elif k == N.TILE_CPHASE:
        need = cm | (1 << op.q0) | (1 << op.q1)
        X = np.where((idx & need) == need, X * complex(mm[0], mm[1]), X)
    elif k == N.TILE_GDIAG:
        gb = gbits or {}
        ng = op.q1 >> 16
        sel = gb.get(op.q1 & 0xFF, 0) | ((gb.get((op.q1 >> 8) & 0xFF, 0) << 1) if ng == 2 else 0)
        d = mm[:16].view(np.complex128)
        X = X * d[2 * sel + ((idx >> op.q0) & 1)]
    else:
        raise AssertionError(k)
    return X


def _emulate_slot_program(sched, m, tb, psi_tile, f32=False, gbits=None):
    """Numpy model of the register tile kernel: X[s] is the amplitude held
    in slot s (register s >> tb of thread s & (2^tb - 1)); ops are in slot
    bits.  The final X must be in the identity layout."""
    X = psi_tile.copy()
    i = 0
    while i < len(sched):
        op = sched[i]
        if op.kind == N.TILE_CUT_PHASE:
            raw = bytes(bytearray(memoryview(op.m).cast("B")))
            E, nd = raw[0], raw[1]
            T = []
            for d in sched[i + 1:i + 1 + nd]:
                assert d.kind == N.TILE_DATA
                mm = np.array([d.m[k] for k in range(32)])
                if f32:
                    mm = np.frombuffer(mm.tobytes(), dtype=np.float32)[:32].astype(np.float64)
                T += list(mm.reshape(-1, 2) @ np.array([1, 1j]))
            X = X * np.array(T)[_cut_counts(raw, m, tb, gbits or {})]
            i += 1 + nd
            continue
        X = _apply_vec(X, op, m, tb, f32, gbits)
        i += 1
    return X


def _cut_counts(raw, m, tb, gbits):
    """Index into the phase table of every slot: the cut count that the
    kernel forms from the CUT_PHASE bytes (pairs by class, masks B[r])."""
    idx = np.arange(1 << m)
    t, r = idx & ((1 << tb) - 1), idx >> tb
    E = raw[0]
    c = np.zeros(idx.size, dtype=np.int64)
    A = np.zeros(idx.size, dtype=np.int64)
    for i in range(E):
        cls, e0, e1 = raw[72 + i], raw[8 + 2 * i], raw[9 + 2 * i]
        if cls == 0:
            c += ((t >> e0) ^ (t >> e1)) & 1
        elif cls == 1:
            c += ((t >> e0) & 1) ^ gbits.get(e1, 0)
        elif cls == 2:
            A |= ((t >> e0) & 1) << i
        elif cls == 3:
            A |= gbits.get(e1, 0) << i
        elif cls == 5:
            c += gbits.get(e0, 0) ^ gbits.get(e1, 0)
        else:
            assert cls == 4
    B = np.frombuffer(raw[128:192], dtype=np.uint32).astype(np.int64)
    x = A ^ B[r]
    pop = np.array([bin(int(v)).count("1") for v in x])
    return c + pop


def _op(kind, q0, q1, ctrl, values):
    op = N.TileOp()
    op.kind, op.q0, op.q1, op.ctrl_mask = kind, q0, q1, ctrl
    v = np.zeros(16, dtype=np.complex128)
    vals = np.asarray(values, dtype=np.complex128).reshape(-1)
    v[:vals.size] = vals
    for i, x in enumerate(v.view(np.float64)):
        op.m[i] = x
    return op


def _rand_1q(rng):
    """A Haar 2x2, a real rotation (RY) or an RX: the last two have zero
    parts, which the scheduler turns into its reduced register kinds."""
    c = rng.integers(3)
    th = rng.uniform(0, 6)
    if c == 1:
        return np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]], dtype=np.complex128)
    if c == 2:
        return np.array([[np.cos(th), -1j * np.sin(th)], [-1j * np.sin(th), np.cos(th)]])
    return np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]


def _random_tile_ops(m, count, rng, outside=()):
    """Random ops of every public kind on an m-qubit tile; GDIAG ops (one
    or two of the state qubits ``outside`` the tile) when outside is given."""
    ops = []
    kinds = [1, 1, 2, 3, 4, 5, 5, 6, 6, 7] + ([8, 8] if len(outside) else []) + [9, 9, 9]
    zz = np.exp(1j * rng.uniform(0, 6, 2))  # one ZZ angle: diag(a, b, b, a)
    for _ in range(count):
        kind = int(rng.choice(kinds))
        q = [int(x) for x in rng.choice(m, 3, replace=False)]
        ctrl = 0
        if rng.random() < 0.2:
            ctrl = 1 << q[2]
        if kind == 9:  # a ZZ phase on a pair of tile bits or of tile and outside qubits
            a, b = zz
            where = int(rng.integers(3)) if len(outside) >= 2 else 0
            if where == 0:
                ops.append(_op(N.TILE_DIAG2, q[0], q[1], 0, np.array([a, b, b, a])))
            elif where == 1:
                g = int(rng.choice(list(outside)))
                ops.append(_op(N.TILE_GDIAG, q[0], g | (1 << 16), 0, np.array([a, b, b, a])))
            else:
                g0, g1 = [int(x) for x in rng.choice(list(outside), 2, replace=False)]
                ops.append(_op(N.TILE_GDIAG, 0, g0 | (g1 << 8) | (2 << 16), 0, np.array([a, a, b, b, b, b, a, a])))
            continue
        if kind == N.TILE_GDIAG:
            ng = 1 if len(outside) == 1 else int(rng.integers(1, 3))
            g = [int(x) for x in rng.choice(list(outside), ng, replace=False)] + [0]
            ops.append(_op(kind, q[0], g[0] | (g[1] << 8 if ng == 2 else 0) | (ng << 16), 0,
                           np.exp(1j * rng.uniform(0, 6, 2 << ng))))
        elif kind in (N.TILE_DENSE1,):
            ops.append(_op(kind, q[0], 0, ctrl, _rand_1q(rng).reshape(-1)))
        elif kind == N.TILE_DENSE2:
            U = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
            ops.append(_op(kind, q[0], q[1], ctrl, U.reshape(-1)))
        elif kind == N.TILE_DIAG1:
            ops.append(_op(kind, q[0], 0, ctrl, np.exp(1j * rng.uniform(0, 6, 2))))
        elif kind == N.TILE_DIAG2:
            ops.append(_op(kind, q[0], q[1], ctrl, np.exp(1j * rng.uniform(0, 6, 4))))
        elif kind == N.TILE_CPHASE:
            ops.append(_op(kind, q[0], q[1], ctrl, np.array([np.exp(1j * rng.uniform(0, 6))])))
        else:  # CPERM / SWAP
            ops.append(_op(kind, q[0], q[1], ctrl if kind == N.TILE_CPERM or rng.random() < 0.5 else 0,
                                 np.zeros(1)))
    return ops


@pytest.mark.parametrize("dtype,m", [("f64", 10), ("f64", 12), ("f32", 12)])
def test_cut_phase_long_runs(dtype, m):
    """QAOA-like passes (ZZ phases of one angle on many pairs, then RX on
    every tile bit, two layers) under FMA scheduling: runs of up to 32 pairs
    become one CUT_PHASE op with its phase table in several DATA ops."""
    rng = np.random.default_rng(40 + m)
    rb = min(m - 6, 4 if dtype == "f32" else 3)
    tb = m - rb
    outside = (m, m + 1, m + 2)
    gbits = {q: int(rng.integers(2)) for q in outside}
    ops = []
    for layer in range(2):
        a, b = np.exp(1j * rng.uniform(0, 6, 2))
        pairs = [(x, y) for x in range(m) for y in range(x + 1, m)]
        rng.shuffle(pairs)
        for x, y in pairs[:30]:
            ops.append(_op(N.TILE_DIAG2, x, y, 0, np.array([a, b, b, a])))
        for x in range(m):
            ops.append(_op(N.TILE_GDIAG, x, outside[x % 3] | (1 << 16), 0, np.array([a, b, b, a])))
        ops.append(_op(N.TILE_GDIAG, 0, outside[0] | (outside[2] << 8) | (2 << 16), 0,
                       np.array([a, a, b, b, b, b, a, a])))
        th = rng.uniform(0, 6)
        for x in range(m):
            ops.append(_op(N.TILE_DENSE1, x, 0, 0, np.array([np.cos(th), -1j * np.sin(th), -1j * np.sin(th),
                                                           np.cos(th)])))
    sched = N.tile_schedule(dtype, m, ops, True)
    cuts = [o for o in sched if o.kind == N.TILE_CUT_PHASE]
    E = [bytes(bytearray(memoryview(o.m).cast("B")))[0] for o in cuts]
    assert max(E) == 32 and len(cuts) >= 4
    v = rng.normal(size=1 << m) + 1j * rng.normal(size=1 << m)
    want = v.copy()
    for op in ops:
        want = _apply_vec(want, op, m, gbits=gbits)
    got = _emulate_slot_program(sched, m, tb, v, dtype == "f32", gbits)
    assert np.max(np.abs(got - want)) < (1e-11 if dtype == "f64" else 1e-3)


@pytest.mark.parametrize("dtype,m,fma", [("f64", 8, False), ("f64", 9, False), ("f64", 11, False), ("f64", 12, True),
                                         ("f32", 10, False), ("f32", 12, True), ("f32", 13, True)])
def test_tile_schedule_matches_op_list(dtype, m, fma):
    rng = np.random.default_rng(m + (100 if dtype == "f32" else 0))
    rb = min(m - 6, 4 if (dtype == "f32" and fma) else 3)
    tb = m - rb
    kinds = set()
    for trial in range(3):
        # qubits m and m + 1 stand for state qubits outside the tile, with
        # fixed values over the block
        ops = _random_tile_ops(m, 40, rng, outside=(m, m + 1))
        gbits = {m: int(rng.integers(2)), m + 1: int(rng.integers(2))}
        sched = N.tile_schedule(dtype, m, ops, fma)
        v = rng.normal(size=1 << m) + 1j * rng.normal(size=1 << m)
        # reference: the op list applied to the tile as an m-qubit state
        want = v.copy()
        for op in ops:
            want = _apply_vec(want, op, m, gbits=gbits)
        if m <= 9:  # the vector model against the Kronecker model of the op semantics
            base = (gbits[m] << m) | (gbits[m + 1] << (m + 1))
            ref = np.zeros(1 << (m + 2), dtype=np.complex128)
            ref[base:base + (1 << m)] = v
            for op in ops:
                ref = emulate_tile_op(op, list(range(m)), m + 2, ref)
            assert np.max(np.abs(ref[base:base + (1 << m)] - want)) < 1e-12
        got = _emulate_slot_program(sched, m, tb, v, dtype == "f32", gbits)
        assert np.max(np.abs(got - want)) < (1e-12 if dtype == "f64" else 1e-4)
        nremap = sum(1 for o in sched if o.kind == N.TILE_REMAP)
        assert nremap <= len(ops)
        kinds |= {o.kind for o in sched}
    assert N.TILE_D1REAL in kinds and N.TILE_D1RX in kinds
    assert (N.TILE_CUT_PHASE in kinds) == fma
    # a pass whose targets are all register bits needs no relayout
    hi = list(range(tb, m))
    ops = [_op(N.TILE_DENSE1, q, 0, 0, np.eye(2).reshape(-1)) for q in hi]
    ops += [_op(N.TILE_DIAG1, 0, 0, 0, np.array([1, 1j]))]
    assert all(o.kind != N.TILE_REMAP for o in N.tile_schedule(dtype, m, ops, fma))


def test_library_load_refuses_a_stale_build(monkeypatch):
    """The library carries the hash of the sources it was built from; a tree
    whose sources hash differently refuses to load it (no silently stale
    binary on a machine that runs the prebuilt library)."""
    from qvm_amd import _native as N
    from qvm_amd import build as B
    assert open(B.STAMP).read().strip() == B.source_hash()
    saved = N._lib
    try:
        N._lib = None
        monkeypatch.setattr(B, "source_hash", lambda: "0" * 64)
        with pytest.raises(ImportError, match="not built from the sources"):
            N.lib()
    finally:
        N._lib = saved
