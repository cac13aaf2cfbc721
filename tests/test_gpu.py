"""GPU parity tests: the HIP engine through its C ABI against the CPU oracle.

FP64 gate application is compared bit-for-bit (the engine's plain kernels use
the reference serial kernels' multiply-then-add arithmetic); reductions to
1e-14 relative; sampling exactly, except draws within the FP64 summation
bound of a CDF boundary (tests/sampling_check.py), which are counted with the
draws within 1e-15 of a boundary.  FP32 and fused paths are checked by
infidelity / max amplitude error against the FP64 oracle.
"""
import math
import sys

import numpy as np
import pytest

from qvm_amd import _native as N
from qvm_amd import circuits as CI
from qvm_amd import gates as G
from qvm_amd import planner as P
from tests.sampling_check import check_draws

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def S():
    from qvm_amd.state import HipPureState
    assert N.device_count() >= 1, "no HIP device visible"
    return HipPureState


def rand_unitary(k, rng):
    Z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    Qm, R = np.linalg.qr(Z)
    return Qm * (np.diag(R) / np.abs(np.diag(R)))


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return np.ascontiguousarray(v / np.linalg.norm(v))


def load(S, psi, dtype="f64"):
    n = int(psi.size).bit_length() - 1
    st = S(n, dtype)
    st.write_amplitudes(psi)
    return st


# --- gates: bit-exact FP64 ---------------------------------------------------

@pytest.mark.parametrize("n", [1, 5, 12, 17])
def test_1q_every_target_bit_exact(S, oracle, n):
    rng = np.random.default_rng(n)
    psi = rand_state(n, rng)
    st = load(S, psi)
    ref = psi.copy()
    for q in range(n):
        U = rand_unitary(1, rng)
        say(f"1q n={n} q={q}")
        st.apply_matrix(U, [q])
        oracle.apply_matrix(ref, U, [q])
    np.testing.assert_array_equal(st.amplitudes(), ref)
    st.close()


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6])
def test_dense_kq_bit_exact(S, oracle, k):
    rng = np.random.default_rng(100 + k)
    n = 14
    psi = rand_state(n, rng)
    st = load(S, psi)
    ref = psi.copy()
    for trial in range(4):
        qs = [int(x) for x in rng.choice(n, size=k, replace=False)]  # unsorted, non-adjacent
        U = rand_unitary(k, rng)
        say(f"dense k={k} qubits={qs}")
        st.apply_matrix(U, qs)
        oracle.apply_matrix(ref, U, qs)
    np.testing.assert_array_equal(st.amplitudes(), ref)
    st.apply_matrix_dense(rand_unitary(k, rng), list(range(k)))
    st.close()


def test_structured_gates_bit_exact(S, oracle):
    rng = np.random.default_rng(7)
    n = 13
    psi = rand_state(n, rng)
    st = load(S, psi)
    ref = psi.copy()
    gates = [("CNOT", [3, 11]), ("CNOT", [11, 0]), ("CZ", [2, 9]), ("SWAP", [12, 1]),
             ("CCNOT", [0, 7, 4]), ("CSWAP", [5, 12, 2]), ("PHASE", [6]), ("RZ", [4]),
             ("RZZ", [10, 3]), ("CPHASE", [1, 8]), ("X", [0]), ("Z", [12]), ("S", [3]), ("ISWAP", [4, 9]),
             ("PSWAP", [7, 2])]
    for name, qs in gates:
        npar = 1 if name in ("PHASE", "RZ", "RZZ", "CPHASE", "PSWAP") else 0
        U = G.gate_matrix(name, [float(rng.uniform(0, 2 * np.pi))] if npar else [])
        say(f"structured {name} {qs}")
        st.apply_matrix(U, G.nat_tuple(*qs))
        oracle.apply_matrix(ref, U, G.nat_tuple(*qs))
    U = G.controlled(rand_unitary(1, rng))  # CONTROLLED U
    st.apply_matrix(U, G.nat_tuple(9, 2))
    oracle.apply_matrix(ref, U, G.nat_tuple(9, 2))
    np.testing.assert_array_equal(st.amplitudes(), ref)
    # diagonal / permutation entry points
    d = np.exp(1j * rng.uniform(0, 6, 8))
    st.apply_diagonal(d, [1, 6, 11])
    oracle.apply_matrix(ref, np.diag(d), [1, 6, 11])
    perm = [int(x) for x in rng.permutation(8)]
    st.apply_permutation(perm, [12, 0, 5])
    oracle.apply_permutation(ref, perm, [12, 0, 5])
    np.testing.assert_array_equal(st.amplitudes(), ref)
    st.close()


def test_layered_n20_bit_exact_and_measure_all_shots(S, oracle):
    """Config 1 (CPU-runnable): n=20, 20 layers, then 10^4 measure-all shots."""
    n = 20
    gates = CI.layered(n, 20, seed=1)
    st = S(n, "f64")
    ref = oracle.zero_state(n)
    say(f"layered n={n} gates={len(gates)}")
    for U, q in gates:
        st.apply_matrix(U, G.nat_tuple(*q))
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    got = st.amplitudes()
    np.testing.assert_array_equal(got, ref)
    u = CI.uniforms(20201, 10000)
    shots, near = st.sample_ex(u, "gt", 1e-15)
    np.testing.assert_array_equal(shots, st.sample_measure_all(u))
    cdf = oracle.cdf(ref)
    want = np.array([oracle.sample_bisect(ref, p) for p in u[:2000]], dtype=np.uint64)
    r = check_draws(cdf, u[:2000], shots[:2000], want, 1e-15, near[:2000])
    say(f"measure-all shots vs oracle: {r}")
    # a single measure-all collapses to |b> with amplitude exactly 1
    b = st.measure_all(float(u[0]))
    assert b == shots[0]
    amps = st.amplitudes()
    assert amps[b] == 1 and np.count_nonzero(amps) == 1
    st.close()


def test_qft_n14(S, oracle):
    n = 14
    st = S(n, "f64")
    st.apply_matrix(G.gate_matrix("X"), [3])
    ref = oracle.zero_state(n)
    oracle.apply_matrix(ref, G.gate_matrix("X"), [3])
    for U, q in CI.qft(n):
        st.apply_matrix(U, G.nat_tuple(*q))
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    got = st.amplitudes()
    np.testing.assert_array_equal(got, ref)
    N_ = 1 << n
    want = np.exp(2j * np.pi * 8 * np.arange(N_) / N_) / math.sqrt(N_)
    assert np.max(np.abs(got - want)) < 1e-12
    st.close()


# --- reductions and measurement -------------------------------------------

def test_reductions(S, oracle):
    rng = np.random.default_rng(11)
    for n in (1, 6, 15, 21):
        psi = rand_state(n, rng)
        st = load(S, psi)
        for q in sorted({0, n // 2, n - 1}):
            say(f"reduce n={n} q={q}")
            assert abs(st.prob_excited(q) - oracle.excited_probability(psi, q)) < 1e-14
            assert abs(st.prob_ground(q) - oracle.ground_probability(psi, q)) < 1e-14
        assert abs(st.norm() - oracle.norm(psi)) < 1e-14
        st.write_amplitudes(psi * 2.5)
        st.normalize()
        assert abs(st.norm() - 1) < 1e-14
        pr = st.probabilities()
        assert np.allclose(pr, np.abs(psi) ** 2, rtol=1e-13, atol=1e-300)
        st.close()


@pytest.mark.parametrize("compiled", [False, True])
def test_measure_matches_oracle(S, oracle, compiled):
    rng = np.random.default_rng(12 + compiled)
    n = 11
    psi = rand_state(n, rng)
    st = load(S, psi)
    ref = psi.copy()
    for q in (0, 4, 10, 4):
        u = float(rng.random())
        say(f"measure q={q} u={u} compiled={compiled}")
        b, used = st.measure(q, u, compiled)
        bo, usedo = oracle.measure(ref, q, u, compiled)
        assert (b, used) == (bo, usedo)
    assert np.max(np.abs(st.amplitudes() - ref)) < 1e-13
    # p1 == 0: no draw (interpreted)
    st0 = S(3, "f64")
    assert st0.measure(2, 0.3, False) == (0, 0)
    assert st0.measure(2, 0.3, True) == (0, 1)
    st0.close()
    st.close()


def test_force_measure_and_errors(S, oracle):
    rng = np.random.default_rng(13)
    psi = rand_state(9, rng)
    st = load(S, psi)
    ref = psi.copy()
    p1 = oracle.excited_probability(psi, 5)
    st.force_measure(5, 1, p1)
    oracle.force_measurement(ref, 5, 1, p1)
    np.testing.assert_array_equal(st.amplitudes(), ref)
    with pytest.raises(N.QVMHipError) as e:
        st.measure(9, 0.5)
    assert e.value.code == N.EINVAL
    with pytest.raises(N.QVMHipError):
        st.apply_matrix(np.eye(2), [12])
    with pytest.raises(N.QVMHipError):
        st.apply_matrix(np.eye(4), [1, 1])
    with pytest.raises(N.QVMHipError) as e:
        st.measure_all(1.5)
    assert e.value.code == N.ERANGE
    st.close()
    # forcing an outcome of probability 0 has no normalisation: an error, and
    # the state is left as it was
    z = S(4, "f64")
    with pytest.raises(N.QVMHipError) as e:
        z.force_measure(2, 1, 0.0)
    assert e.value.code == N.ERANGE
    want = np.zeros(16, complex)
    want[0] = 1
    np.testing.assert_array_equal(z.amplitudes(), want)
    # compiled projector at the end of the draw range: p0 = 1 keeps 0 for u = 1
    assert z.measure(2, 1.0, True) == (0, 1)
    np.testing.assert_array_equal(z.amplitudes(), want)
    z.close()


def test_sample_cdf_rule(S, oracle):
    rng = np.random.default_rng(14)
    for n in (1, 3, 8, 13, 18, 22):
        psi = rand_state(n, rng)
        psi[rng.integers(0, psi.size, psi.size // 4)] = 0  # zero-probability runs
        psi /= np.linalg.norm(psi)
        st = load(S, psi)
        cdf = oracle.cdf(psi)
        u = np.concatenate([rng.random(3000), cdf[: min(64, cdf.size)], [0.0, 1.0]])
        say(f"sample n={n} draws={u.size}")
        got, near = st.sample_ex(u, "ge", 1e-15)
        np.testing.assert_array_equal(got, st.sample(u))
        want = oracle.sample_multiple(psi, u)
        r = check_draws(cdf, u, got, want, 1e-15, near)
        say(f"  cdf >= u: {r}")
        got2, near2 = st.sample_ex(u, "gt", 1e-15)
        np.testing.assert_array_equal(got2, st.sample_measure_all(u))
        want2 = np.array([oracle.sample_bisect(psi, p) for p in u[:500]], dtype=np.uint64)
        r2 = check_draws(cdf, u[:500], got2[:500], want2, 1e-15, near2[:500])
        say(f"  C > u: {r2}")
        st.close()


@pytest.mark.parametrize("n", [6, 12, 20])
def test_sample_dyadic_ties_exact(S, oracle, n):
    """Dyadic probabilities whose partial sums are exact in every order:
    draws placed exactly on CDF values must resolve the tie by each rule
    (C >= u picks the index whose C equals u, C > u the next nonzero one)
    with no mismatch, and the engine must flag each as near a boundary."""
    rng = np.random.default_rng(40 + n)
    m = rng.integers(0, 32, 1 << n).astype(np.float64)
    m[0] = 1.0
    # amplitudes m * 2^-e: the squares m^2 * 2^-2e and every partial sum of
    # them are exact doubles (fewer than 2^30 units of 2^-2e in all)
    e = (int(np.sum(m * m)).bit_length() + 1) // 2
    psi = (m * 2.0 ** -e).astype(np.complex128)
    prob = (m * m) * 2.0 ** (-2 * e)
    st = load(S, psi)
    cdf = np.cumsum(prob)
    pick = np.unique(rng.integers(0, psi.size, 400))
    u = np.concatenate([cdf[pick], rng.random(1000)])
    u = u[u <= 1.0]
    got, near = st.sample_ex(u, "ge", 1e-15)
    want = oracle.sample_multiple(psi, u)
    np.testing.assert_array_equal(got, want)
    assert near[: pick.size].all()
    got2, near2 = st.sample_ex(u, "gt", 1e-15)
    want2 = np.array([oracle.sample_bisect(psi, p) for p in u], dtype=np.uint64)
    np.testing.assert_array_equal(got2, want2)
    assert near2[: pick.size].all()
    # the same draws with eps = 0 flag nothing
    _, none = st.sample_ex(u, "ge", 0.0)
    assert not none.any()
    st.close()


def test_sample_goldens(S, oracle):
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
    for case in gold["sample_bisect"]["cases"]:
        psi = np.array(case["amplitudes"], dtype=np.complex128)
        st = load(S, psi)
        u = np.array([p for p, _ in case["draws"]])
        want = [w for _, w in case["draws"]]
        assert list(st.sample_measure_all(u)) == want
        st.close()


# --- amplitudes, copies, expectations ---------------------------------------

def test_amplitude_io_copy_inner_expectation(S):
    rng = np.random.default_rng(15)
    n = 12
    psi = rand_state(n, rng)
    a = load(S, psi)
    np.testing.assert_array_equal(a.amplitudes(), psi)
    np.testing.assert_array_equal(a.read_amplitudes(100, 7), psi[100:107])
    b = S(n, "f64")
    b.copy_from(a)
    np.testing.assert_array_equal(b.amplitudes(), psi)
    phi = rand_state(n, rng)
    b.write_amplitudes(phi)
    ip = a.inner_product(b)
    assert abs(ip - np.vdot(psi, phi)) < 1e-13
    edges = [(0, 3), (3, 7), (7, 11), (2, 5)]
    idx = np.arange(psi.size)
    zz = sum(np.sum(np.abs(psi) ** 2 * (1 - 2 * (((idx >> x) ^ (idx >> y)) & 1))) for x, y in edges)
    assert abs(a.expectation_zz(edges) - zz) < 1e-13
    a.reset_zero()
    z = a.amplitudes()
    assert z[0] == 1 and np.count_nonzero(z) == 1
    a.close()
    b.close()


def test_copy_and_accumulate_read_src_before_its_next_call(S):
    """copy / accumulate run on the destination's stream; the source's next
    call (on its own stream) must not change it before they have read it."""
    rng = np.random.default_rng(16)
    n = 22
    psi = rand_state(n, rng)
    src = load(S, psi)
    dst = S(n, "f64")
    acc = S(n, "f64")
    acc.write_amplitudes(psi)
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    for rep in range(3):
        say(f"copy/accumulate ordering rep={rep}")
        snap = src.amplitudes()
        dst.copy_from(src)
        acc.accumulate(src)
        for q in range(n):  # queued on src's stream right behind the reads
            src.apply_matrix(H, [q])
        np.testing.assert_array_equal(dst.amplitudes(), snap)
        np.testing.assert_array_equal(acc.amplitudes(), (psi if rep == 0 else prev_acc) + snap)
        prev_acc = acc.amplitudes()
    for st in (src, dst, acc):
        st.close()


def test_determinism(S):
    n = 16
    gates = CI.haar_cz(n, 3, seed=9)
    outs = []
    for _ in range(2):
        st = S(n, "f64")
        for U, q in gates:
            st.apply_matrix(U, G.nat_tuple(*q))
        outs.append((st.amplitudes(), st.prob_excited(3), st.norm()))
        st.close()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1] and outs[0][2] == outs[1][2]


# --- tile passes, fused blocks, FP32, programs ------------------------------

@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("m", [8, 9, 10, 11, 12])
def test_tile_passes_vs_plain(S, oracle, m, reorder):
    """In-order tile passes reproduce the plain path bit for bit; reordered
    passes (commuting gates moved) agree to rounding."""
    n = 16
    gates = CI.layered(n, 6, seed=5) + CI.qft(n)[:40] + [(G.gate_matrix("RZZ", [0.7]), [3, 14]),
                                                         (G.controlled(rand_unitary(1, np.random.default_rng(1))), [15, 2]),
                                                         (G.gate_matrix("CCNOT"), [1, 9, 13])]
    plan = P.plan_tiled(gates, n, m=m, low=4, reorder=reorder)
    say(f"tile m={m} reorder={reorder} passes={sum(1 for it in plan if it.kind == 'tile')} items={len(plan)}")
    st = S(n, "f64")
    P.execute_plan(st, plan)
    ref = oracle.zero_state(n)
    for U, q in gates:
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    got = st.amplitudes()
    if reorder:
        assert np.max(np.abs(got - ref)) < 1e-14
    else:
        np.testing.assert_array_equal(got, ref)
    st.close()


def _random_tile_ops(m, count, rng, outside=()):
    """Random tile ops of every public kind, some controlled (local bits);
    GDIAG ops on one or two of the state qubits ``outside`` the tile."""
    ops = []
    kinds = [1, 1, 2, 3, 4, 5, 5, 6, 6, 7] + ([8, 8] if len(outside) else []) + [9, 9, 9]
    za, zb = np.exp(1j * rng.uniform(0, 6, 2))  # one ZZ angle, diag(a, b, b, a)
    for _ in range(count):
        kind = int(rng.choice(kinds))
        q = [int(x) for x in rng.choice(m, 3, replace=False)]
        ctrl = (1 << q[2]) if rng.random() < 0.25 else 0
        op = N.TileOp()
        if kind == 9:  # ZZ phases (FMA passes merge runs of them into one cut-phase op)
            where = int(rng.integers(3)) if len(outside) >= 2 else 0
            op.ctrl_mask = 0
            if where == 0:
                op.kind, op.q0, op.q1 = N.TILE_DIAG2, q[0], q[1]
                vals = np.array([za, zb, zb, za])
            elif where == 1:
                op.kind, op.q0, op.q1 = N.TILE_GDIAG, q[0], int(rng.choice(list(outside))) | (1 << 16)
                vals = np.array([za, zb, zb, za])
            else:
                g0, g1 = [int(x) for x in rng.choice(list(outside), 2, replace=False)]
                op.kind, op.q0, op.q1 = N.TILE_GDIAG, 0, g0 | (g1 << 8) | (2 << 16)
                vals = np.array([za, za, zb, zb, zb, zb, za, za])
            v = np.zeros(16, dtype=np.complex128)
            v[:vals.size] = vals
            for k, x in enumerate(v.view(np.float64)):
                op.m[k] = x
            ops.append(op)
            continue
        op.kind, op.q0, op.ctrl_mask = kind, q[0], ctrl
        op.q1 = q[1] if kind not in (N.TILE_DENSE1, N.TILE_DIAG1) else 0
        if kind == N.TILE_GDIAG:
            ng = 1 if len(outside) == 1 else int(rng.integers(1, 3))
            g = [int(x) for x in rng.choice(list(outside), ng, replace=False)] + [0]
            op.ctrl_mask = 0
            op.q1 = g[0] | ((g[1] << 8) if ng == 2 else 0) | (ng << 16)
            vals = np.exp(1j * rng.uniform(0, 6, 2 << ng))
        elif kind == N.TILE_DENSE1:
            c, th = rng.integers(3), rng.uniform(0, 6)
            if c == 1:  # real rotation and RX: the scheduler's reduced kinds
                vals = np.array([np.cos(th), -np.sin(th), np.sin(th), np.cos(th)], dtype=np.complex128)
            elif c == 2:
                vals = np.array([np.cos(th), -1j * np.sin(th), -1j * np.sin(th), np.cos(th)])
            else:
                vals = rand_unitary(1, rng).reshape(-1)
        elif kind == N.TILE_DENSE2:
            vals = rand_unitary(2, rng).reshape(-1)
        elif kind == N.TILE_DIAG1:
            vals = np.exp(1j * rng.uniform(0, 6, 2))
        elif kind == N.TILE_DIAG2:
            vals = np.exp(1j * rng.uniform(0, 6, 4))
        elif kind == N.TILE_CPHASE:
            vals = np.exp(1j * rng.uniform(0, 6, 1))
        else:
            vals = np.zeros(1)
            if kind == N.TILE_SWAP and rng.random() < 0.5:
                op.ctrl_mask = 0
        v = np.zeros(16, dtype=np.complex128)
        v[:vals.size] = vals
        for k, x in enumerate(v.view(np.float64)):
            op.m[k] = x
        ops.append(op)
    return ops


def _tile_op_as_gate(op, tq):
    """(U, Quil-order qubits) of a tile op on tile qubits tq."""
    mm = np.array([op.m[i] for i in range(32)]).view(np.complex128)
    g = lambda b: tq[b]  # noqa: E731
    if op.kind == N.TILE_DENSE1:
        U, qs = mm[:4].reshape(2, 2), [g(op.q0)]
    elif op.kind == N.TILE_DIAG1:
        U, qs = np.diag(mm[:2]), [g(op.q0)]
    elif op.kind == N.TILE_DENSE2:
        U, qs = mm[:16].reshape(4, 4), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_DIAG2:
        U, qs = np.diag(mm[:4]), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_CPERM:
        U, qs = G.gate_matrix("CNOT"), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_SWAP:
        U, qs = G.gate_matrix("SWAP"), [g(op.q1), g(op.q0)]
    elif op.kind == N.TILE_GDIAG:
        ng = op.q1 >> 16
        outside = [op.q1 & 0xFF, (op.q1 >> 8) & 0xFF][:ng]
        U, qs = np.diag(mm[:1 << (1 + ng)]), list(reversed([g(op.q0)] + outside))
    else:
        U, qs = np.diag([1, 1, 1, mm[0]]), [g(op.q1), g(op.q0)]
    for c in [b for b in range(32) if (op.ctrl_mask >> b) & 1]:
        U, qs = G.controlled(U), [g(c)] + qs
    return U, qs


@pytest.mark.parametrize("dtype,m,fma", [("f64", 8, False), ("f64", 10, False), ("f64", 12, False),
                                         ("f64", 11, True), ("f32", 9, False), ("f32", 13, True)])
def test_tile_pass_every_op_kind(S, oracle, dtype, m, fma):
    """Random passes of every op kind (controlled dense 1q/2q, diagonals on
    thread and register bits, CNOT, controlled and plain SWAP, CPHASE,
    diagonals on qubits outside the tile) through
    qvmhip_apply_tile_pass: exact FP64 passes equal the oracle bit for bit."""
    from qvm_amd.state import HipPureState
    rng = np.random.default_rng(m * 7 + (1 if fma else 0))
    n = 15
    tq = sorted([0, 1, 2] + [int(x) for x in rng.choice(np.arange(3, n), m - 3, replace=False)])
    psi = rand_state(n, rng)
    st = HipPureState(n, dtype, fma_tiles=fma)
    st.write_amplitudes(psi)
    ref = psi.copy()
    for rep in range(3):
        ops = _random_tile_ops(m, 50, rng, outside=[q for q in range(n) if q not in tq])
        say(f"tile ops {dtype} m={m} fma={fma} rep={rep}")
        st.apply_tile_pass(tq, ops)
        for op in ops:
            U, qs = _tile_op_as_gate(op, tq)
            oracle.apply_matrix(ref, U, G.nat_tuple(*qs))
    got = st.amplitudes()
    if dtype == "f64" and not fma:
        np.testing.assert_array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) < (1e-13 if dtype == "f64" else 2e-6)
    st.close()


@pytest.mark.parametrize("dtype,m", [("f64", 10), ("f64", 12), ("f32", 12)])
def test_tile_pass_long_zz_runs(S, oracle, dtype, m):
    """QAOA-shaped FMA passes: runs of up to 32 ZZ phases of one angle
    (pairs of tile qubits, tile and outside qubits, two outside qubits) and
    RX on every tile qubit, against the oracle applying each gate."""
    from qvm_amd.state import HipPureState
    rng = np.random.default_rng(70 + m)
    n = 15
    tq = sorted([0, 1, 2] + [int(x) for x in rng.choice(np.arange(3, n), m - 3, replace=False)])
    outside = [q for q in range(n) if q not in tq]
    psi = rand_state(n, rng)
    st = HipPureState(n, dtype, fma_tiles=True)
    st.write_amplitudes(psi)
    ref = psi.copy()
    ops = []

    def op_of(kind, q0, q1, vals):
        op = N.TileOp()
        op.kind, op.q0, op.q1, op.ctrl_mask = kind, q0, q1, 0
        v = np.zeros(16, dtype=np.complex128)
        v[:len(vals)] = vals
        for k, x in enumerate(v.view(np.float64)):
            op.m[k] = x
        return op
    for layer in range(2):
        a, b = np.exp(1j * rng.uniform(0, 6, 2))
        pairs = [(x, y) for x in range(m) for y in range(x + 1, m)]
        rng.shuffle(pairs)
        ops += [op_of(N.TILE_DIAG2, x, y, [a, b, b, a]) for x, y in pairs[:30]]
        ops += [op_of(N.TILE_GDIAG, x, outside[x % len(outside)] | (1 << 16), [a, b, b, a]) for x in range(m)]
        ops.append(op_of(N.TILE_GDIAG, 0, outside[0] | (outside[1] << 8) | (2 << 16), [a, a, b, b, b, b, a, a]))
        th = rng.uniform(0, 6)
        ops += [op_of(N.TILE_DENSE1, x, 0, [np.cos(th), -1j * np.sin(th), -1j * np.sin(th), np.cos(th)])
                for x in range(m)]
    say(f"long zz runs {dtype} m={m} ops={len(ops)}")
    st.apply_tile_pass(tq, ops)
    for op in ops:
        U, qs = _tile_op_as_gate(op, tq)
        oracle.apply_matrix(ref, U, G.nat_tuple(*qs))
    got = st.amplitudes()
    assert np.max(np.abs(got - ref)) < (1e-13 if dtype == "f64" else 2e-6)
    st.close()


def test_tile_pass_f32(S, oracle):
    n = 14
    gates = CI.layered(n, 4, seed=6)
    plan = P.plan_tiled(gates, n, m=13, low=5, merge_1q=True)
    st = S(n, "f32")
    P.execute_plan(st, plan)
    ref = oracle.zero_state(n)
    for U, q in gates:
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    got = st.amplitudes()
    assert np.max(np.abs(got - ref)) < 1e-6
    st.close()


@pytest.mark.parametrize("dtype,m", [("f64", 9), ("f64", 11), ("f64", 12), ("f32", 10), ("f32", 13)])
def test_tile_passes_fma(S, oracle, dtype, m):
    """Tile passes with FMA-contracted products (the fast mode): same state
    as gate-by-gate application within rounding (FP64 1e-13, FP32 2e-6)."""
    from qvm_amd.state import HipPureState
    n = 15
    gates = CI.layered(n, 5, seed=m) + CI.qft(n)[:60] + CI.haar_cz(n, 2, seed=m)
    plan = P.plan_tiled(gates, n, m=m, low=4, merge_1q=True)
    say(f"fma tiles {dtype} m={m} passes={len(plan)}")
    st = HipPureState(n, dtype, fma_tiles=True)
    P.execute_plan(st, plan)
    ref = oracle.zero_state(n)
    for U, q in gates:
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    assert np.max(np.abs(st.amplitudes() - ref)) < (1e-13 if dtype == "f64" else 2e-6)
    st.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("k", [3, 4, 5, 6])
def test_fused_mfma_block(S, oracle, k, dtype):
    """Both column layouts: targets 0..k-1 or n = k + 5 (32 complement
    groups) take the MFMA operand layout (16 groups per wave); targets above
    the line at n = 13 take one row of 64 groups per memory instruction for
    k = 4 and FP32 k = 5."""
    rng = np.random.default_rng(200 + k)
    for n in (k + 5, 13):
        psi = rand_state(n, rng)
        high = [int(x) for x in 4 + rng.choice(n - 4, size=k, replace=False)]
        for qs in (list(range(k)), [int(x) for x in rng.choice(n, size=k, replace=False)], high):
            U = rand_unitary(k, rng)
            say(f"fused k={k} dtype={dtype} n={n} qubits={qs}")
            st = load(S, psi, dtype)
            st.apply_fused(U, qs)
            ref = oracle.apply_matrix(psi.copy(), U, qs)
            err = np.max(np.abs(st.amplitudes() - ref))
            assert err < (1e-14 if dtype == "f64" else 3e-6), err
            st.close()


def test_fused_plan_random_mix(S, oracle):
    n = 16
    gates = CI.random_mix(n, 8, seed=21)
    ref = oracle.zero_state(n)
    for U, q in gates:
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    for kmax in (3, 4, 5):
        plan = P.plan_fused(gates, n, kmax)
        st = S(n, "f64")
        P.execute_plan(st, plan)
        err = np.max(np.abs(st.amplitudes() - ref))
        say(f"fused plan kmax={kmax} blocks={len(plan)} err={err:.2e}")
        assert err < 1e-12
        st.close()


def test_fma_tiled_and_fused_plans_n24_vs_oracle(S, oracle):
    """The reordered FMA tile passes and the fused MFMA blocks against the
    oracle (not only a round trip) at n = 24, where tiles name qubits up to
    23 and blocks span far-apart qubits (k = 4 blocks above the line take the
    row layout)."""
    n = 24
    gates = CI.random_mix(n, 6, seed=24)
    ref = oracle.zero_state(n)
    for U, q in gates:
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
    plans = {"tiled_fma": P.plan_tiled(gates, n, 11, 4, merge_1q=True),
             "tiled_fma_l3": P.plan_tiled(gates, n, 11, 3, merge_1q=True),
             "hybrid": P.plan_tiled(gates, n, 11, 4, merge_1q=True, wide="fused"),
             "fused_k4": P.plan_fused(gates, n, 4), "fused_k5": P.plan_fused(gates, n, 5),
             "fused_k6": P.plan_fused(gates, n, 6)}
    for name, plan in plans.items():
        st = S(n, "f64", fma_tiles=True)
        prog = P.build_program(st, plan, use_graph=True)
        prog.run()
        err = float(np.max(np.abs(st.amplitudes() - ref)))
        say(f"{name} n={n} items={len(plan)} max|d psi|={err:.2e}")
        assert err < 1e-12, name
        prog.close()
        st.close()


def test_f32_layered_infidelity(S, oracle):
    n = 18
    gates = CI.layered(n, 10, seed=3)
    ref = oracle.zero_state(n)
    st = S(n, "f32")
    for U, q in gates:
        oracle.apply_matrix(ref, U, G.nat_tuple(*q))
        st.apply_matrix(U, G.nat_tuple(*q))
    got = st.amplitudes()
    # fidelity of the normalised states: FP32 gate matrices are unitary only
    # to ~6e-8, so the FP32 norm drifts; it is renormalised in FP64.
    infid = 1 - abs(np.vdot(ref, got)) ** 2 / np.vdot(got, got).real
    say(f"f32 infidelity {infid:.3e} max err {np.max(np.abs(got - ref)):.3e} norm {st.norm():.9f}")
    assert infid < 1e-9
    assert abs(st.norm() - 1) < 1e-4
    st.normalize()
    assert abs(st.norm() - 1) < 1e-6
    st.close()


@pytest.mark.parametrize("mode", ["plain", "tiled", "fused"])
def test_program_graph_replay(S, oracle, mode):
    n = 14
    gates = CI.layered(n, 3, seed=8)
    st = S(n, "f64")
    eager = S(n, "f64")
    plan = {"plain": P.plan_plain, "tiled": lambda g, n: P.plan_tiled(g, n, 10, 4),
            "fused": lambda g, n: P.plan_fused(g, n, 4)}[mode](gates, n)
    prog = P.build_program(st, plan, use_graph=True)
    for rep in range(3):
        say(f"graph {mode} replay {rep}")
        prog.run()
        P.execute_plan(eager, plan)
        eager.sync()
        _ = eager.norm()  # eager work between replays
        a, b = st.amplitudes(), eager.amplitudes()
        if mode == "fused":
            assert np.max(np.abs(a - b)) < 1e-13
        else:
            np.testing.assert_array_equal(a, b)
    prog.close()
    st.close()
    eager.close()


def test_path_sum_and_xy_piswap_programs(S, oracle):
    """The reference's MADPI programs (tests/path-simulate-tests.lisp:21-30)
    run on the engine equal their path sums, and XY(pi/2) 1 0 equals
    PISWAP(pi/2) 1 0 on random 2-qubit states (tests/gate-tests.lisp:115-127),
    both also against the oracle."""
    import json
    import os

    from qvm_amd import qvm as QV
    from qvm_amd import quil as Q

    from tests.test_oracle import path_sum

    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
    for lines in gold["path_sum_programs"]["programs"]:
        n = 1 + max(q for ins in Q.parse_quil("\n".join(lines)).executable_code for q in ins.qubits)
        q = QV.run_program(n, "\n".join(lines))
        err = np.max(np.abs(q.amplitudes() - path_sum(lines, n)))
        say(f"path sum {lines}: {err:.1e}")
        assert err < 1e-15, lines
        q.close()
    rng = np.random.default_rng(71)
    xy, pis = ("\n".join(p) for p in gold["xy_piswap"]["programs"])
    for _ in range(10):
        psi = rand_state(2, rng)
        outs = []
        for prog in (xy, pis):
            q = QV.make_qvm(2)
            q.set_amplitudes(psi)
            q.load_program(prog)
            q.run()
            outs.append(q.amplitudes())
            q.close()
            p = Q.parse_quil(prog)
            ref = psi.copy()
            for ins in p.executable_code:
                oracle.apply_matrix(ref, Q.resolve_gate_matrix(ins, p.gate_definitions, {}), G.nat_tuple(*ins.qubits))
            assert np.max(np.abs(outs[-1] - ref)) < 1e-15
        assert np.max(np.abs(outs[0] - outs[1])) < 1e-15


def test_qvm_api_end_to_end(S):
    import json
    import os

    from qvm_amd import qvm as QV
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
    for key in ("simple_measurements", "measurements_with_swap"):
        g = gold[key]
        q = QV.run_program(3, "\n".join(g["program"]))
        assert [int(q.memory_ref("ro", i)) for i in range(3)] == g["ro"]
        q.close()
    for i, lines in enumerate(gold["measure_all_basis"]["programs"]):
        q = QV.make_qvm(3)
        q.load_program("\n".join(lines) or "I 0")
        q.run()
        _, bits = q.measure_all()
        assert bits == [(i >> b) & 1 for b in range(3)]
        assert q.nth_amplitude(i) == 1
        q.close()
    q = QV.run_program(3, "\n".join(gold["dagger_inversion"]["program"]))
    a = q.amplitudes()
    assert abs(a[0] - 1) < 1e-14 and np.max(np.abs(a[1:])) < 1e-14
    q.close()
    # unit norm after measurements (tests/measurement-tests.lisp:88-109).
    # The draws are seeded and replayed on the oracle, so the collapsed state
    # is checked amplitude by amplitude.  The norm drifts from 1 by a few ulps
    # per renormalization, by up to 4e-14 after 9 steps on the oracle itself
    # depending on the outcomes; the reference checks it with its relative
    # epsilon 1e-5 (tests/utilities.lisp:71,103-122).
    from oracle import oracle as O
    from qvm_amd import gates as GT
    for n, seed in ((5, 1), (9, 2), (9, 3), (15, 4)):
        prog = "\n".join(f"H {k}\nMEASURE {k}" for k in range(n))
        draws = np.random.default_rng(seed).random(n)
        it = iter(draws.tolist())
        q = QV.run_program(n, prog, rng=lambda: next(it))
        ref = O.zero_state(n)
        H = GT.gate_matrix("H", [])
        for k in range(n):
            O.apply_matrix(ref, H, [k])
            O.measure(ref, k, float(draws[k]), False)
        a = q.amplitudes()
        assert np.max(np.abs(a - ref)) < 1e-13
        assert abs(np.linalg.norm(a) - 1) < 1e-5
        q.close()
    # compiled program path
    q = QV.make_qvm(10)
    q.load_program("\n".join(f"H {k}" for k in range(10)) + "\nCNOT 0 9\nRZ(0.3) 4")
    q.compile_loaded_program(mode="tiled")
    q.run()
    pr = np.abs(q.amplitudes()) ** 2
    assert np.allclose(pr, 2.0 ** -10)
    q.close()
