"""Plain-kernel dispatch paths against the oracle, one case per path: 1-qubit
dense gates on every target (one amplitude per thread with lane shuffles for
targets < 6, LDS exchange from 6 up in FP64; FP32 and controlled gates take
the pair form), full diagonals of width 1..4 (compile-time entry
selects for 1..3, indexed entries for 4; two amplitudes per access in FP32),
single transpositions whose bits are all low (whole-state lane sweep) or
not, and whole-line sweeps for single-entry diagonals and transpositions with
a bit inside a 128 B line.  FP64 results are bit-identical to the oracle's serial arithmetic; FP32
within 1e-6."""
import sys

import numpy as np
import pytest

from qvm_amd import _native as N
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


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def rand_u(k, rng):
    Z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    return np.linalg.qr(Z)[0]


def check(st, ref, dtype):
    got = st.amplitudes()
    if dtype == "f64":
        np.testing.assert_array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) < 1e-6


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("n", [8, 13])
def test_dense1_every_target(S, oracle, dtype, n):
    rng = np.random.default_rng(n)
    say(f"dense 1q every target n={n} {dtype}")
    psi = rand_state(n, rng)
    st = S(n, dtype)
    st.write_amplitudes(psi)
    ref = psi.astype(np.complex64).astype(np.complex128) if dtype == "f32" else psi.copy()
    for q in list(range(n)) + list(range(n - 1, -1, -1)):
        U = rand_u(1, rng)
        st.apply_matrix(U, [q])
        oracle.apply_matrix(ref, U, [q])
    # controlled 1q gates (pair form) on low targets
    for t, c in ((0, 1), (5, 2), (6, 7), (7, 0), (2, n - 1)):
        U = G.controlled(rand_u(1, rng))
        st.apply_matrix(U, G.nat_tuple(c, t))
        oracle.apply_matrix(ref, U, G.nat_tuple(c, t))
    check(st, ref, dtype)
    st.close()


def f32_apply_1q(psi, U, q):
    """The serial 1q kernel's arithmetic (src/serial-kernels.lisp:19-168: each
    output row u0*a + u1*b, every complex product as (ar*br - ai*bi,
    ar*bi + ai*br), no fused multiply-add) evaluated in float32 with numpy,
    one rounding per operation."""
    x = psi.reshape(-1, 2, 1 << q)
    a, b = x[:, 0, :], x[:, 1, :]
    ur = U.real.astype(np.float32)
    ui = U.imag.astype(np.float32)
    out = np.empty_like(x)
    for r in (0, 1):
        yr = a.real * ur[r, 0] - a.imag * ui[r, 0]
        yi = a.real * ui[r, 0] + a.imag * ur[r, 0]
        pr = b.real * ur[r, 1] - b.imag * ui[r, 1]
        pi = b.real * ui[r, 1] + b.imag * ur[r, 1]
        out[:, r, :].real = yr + pr
        out[:, r, :].imag = yi + pi
    return out.reshape(-1)


@pytest.mark.parametrize("n", [9, 12])
def test_fp32_1q_every_target_bit_exact(S, n):
    """FP32 1q gates, one at a time on every target (two amplitudes per lane
    below target 7 from n = 9, the pair form above), against the serial
    kernel's arithmetic in float32: bit-exact, so the lane pairing and the
    row choice are pinned, not just the value to a tolerance."""
    rng = np.random.default_rng(70 + n)
    say(f"fp32 1q bit-exact n={n}")
    psi = rand_state(n, rng).astype(np.complex64)
    st = S(n, "f32")
    st.write_amplitudes(psi.astype(np.complex128))
    ref = psi.copy()
    for q in list(range(n)) + [0, 6, 3]:
        U = rand_u(1, rng)
        st.apply_matrix(U, [q])
        ref = f32_apply_1q(ref, U, q)
        np.testing.assert_array_equal(np.asarray(st.amplitudes(), dtype=np.complex64), ref)
    st.close()


def test_fp32_single_entry_diagonals_bit_exact(S):
    """FP32 CZ / CPHASE / controlled CPHASE: one changed entry, the amplitudes whose
    listed bits are all 1 multiplied by the entry (serial-kernel products,
    evaluated in float32 here).  Bits all >= 4 take the 16-byte pair kernel,
    lower ones the per-amplitude form; bit-exact for both."""
    n = 14
    rng = np.random.default_rng(83)
    say("fp32 single-entry diagonals bit-exact")
    psi = rand_state(n, rng).astype(np.complex64)
    st = S(n, "f32")
    st.write_amplitudes(psi.astype(np.complex128))
    ref = psi.copy()
    idx = np.arange(1 << n, dtype=np.int64)
    ccphase = G.controlled(G.gate_matrix("CPHASE", [0.4]))
    cases = [(G.gate_matrix("CPHASE", [0.7]), [9, 4]), (G.gate_matrix("CZ"), [13, 5]),
             (G.gate_matrix("CPHASE", [1.9]), [4, 12]), (ccphase, [6, 11, 8]),
             (G.gate_matrix("CPHASE", [2.3]), [9, 1]), (G.gate_matrix("CZ"), [0, 7]),
             (G.gate_matrix("CPHASE", [0.2]), [3, 2])]
    for U, qs in cases:
        st.apply_matrix(U, G.nat_tuple(*qs))
        d = np.complex64(U[-1, -1])
        mask = np.ones(1 << n, dtype=bool)
        for q in qs:
            mask &= ((idx >> q) & 1) == 1
        x = ref[mask]
        dr, di = np.float32(d.real), np.float32(d.imag)
        y = np.empty_like(x)
        y.real = x.real * dr - x.imag * di
        y.imag = x.real * di + x.imag * dr
        ref[mask] = y
        np.testing.assert_array_equal(np.asarray(st.amplitudes(), dtype=np.complex64), ref)
    st.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_full_diagonals_of_every_width(S, oracle, dtype):
    n = 12
    rng = np.random.default_rng(3)
    say(f"full diagonals {dtype}")
    psi = rand_state(n, rng)
    st = S(n, dtype)
    st.write_amplitudes(psi)
    ref = psi.astype(np.complex64).astype(np.complex128) if dtype == "f32" else psi.copy()
    for qs in ([0], [7], [11], [1, 9], [10, 2], [0, 5, 11], [3, 1, 8], [0, 4, 7, 10]):
        k = len(qs)
        D = np.diag(np.exp(1j * rng.uniform(0.1, 6.0, 1 << k)))  # every entry != 1
        st.apply_matrix(D, list(qs))
        oracle.apply_matrix(ref, D, list(qs))
    check(st, ref, dtype)
    st.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_transpositions_low_and_high_bits(S, oracle, dtype):
    n = 11
    rng = np.random.default_rng(5)
    say(f"transpositions {dtype}")
    psi = rand_state(n, rng)
    st = S(n, dtype)
    st.write_amplitudes(psi)
    ref = psi.astype(np.complex64).astype(np.complex128) if dtype == "f32" else psi.copy()
    cases = [("CNOT", [1, 0]), ("CNOT", [0, 2]), ("CNOT", [4, 1]), ("SWAP", [0, 5]), ("SWAP", [2, 1]),
             ("CCNOT", [0, 1, 2]), ("CSWAP", [2, 0, 1]),        # every bit < 6, lowest < 3: lane sweep
             ("CNOT", [6, 0]), ("SWAP", [6, 3]), ("CCNOT", [6, 0, 5]), ("CSWAP", [1, 6, 2]),  # FP32: bits < 7
             ("CNOT", [3, 4]), ("CNOT", [9, 0]), ("SWAP", [1, 10]), ("CCNOT", [8, 2, 5])]
    for name, qs in cases:
        U = G.gate_matrix(name)
        st.apply_matrix(U, G.nat_tuple(*qs))
        oracle.apply_matrix(ref, U, G.nat_tuple(*qs))
    # transpositions only move amplitudes: bit-exact in both dtypes
    np.testing.assert_array_equal(np.asarray(st.amplitudes(), dtype=np.complex128), ref)
    st.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_line_sweeps_for_bits_inside_a_line(S, oracle, dtype):
    """FP64 single-entry diagonals and single transpositions with a bit
    inside a 128 B line (below bit 3) and one above bit 5 sweep whole lines
    (k_diag_line, k_perm_line): entries in two line sets (SWAP, CNOT with the
    control inside the line) or in one (CNOT with the target inside), controls
    inside and outside, every bit inside the line (PHASE, CPHASE 1 0).  FP32
    keeps the entry and pair forms for the same gates."""
    n = 13
    rng = np.random.default_rng(61)
    say(f"line sweeps {dtype}")
    psi = rand_state(n, rng)
    st = S(n, dtype)
    st.write_amplitudes(psi)
    ref = psi.astype(np.complex64).astype(np.complex128) if dtype == "f32" else psi.copy()
    ccp = np.eye(8, dtype=np.complex128)
    ccp[7, 7] = np.exp(0.7j)
    cases = [(G.gate_matrix("CZ"), [0, 9]), (G.gate_matrix("CPHASE", [0.3]), [2, 12]),
             (G.gate_matrix("CPHASE", [1.1]), [1, 0]), (G.gate_matrix("PHASE", [0.9]), [1]),
             (G.gate_matrix("CZ"), [3, 9]), (ccp, [1, 7, 11]), (G.gate_matrix("CPHASE", [0.5]), [11, 0]),
             (G.gate_matrix("SWAP"), [0, 12]), (G.gate_matrix("SWAP"), [2, 7]), (G.gate_matrix("CNOT"), [1, 11]),
             (G.gate_matrix("CNOT"), [10, 2]), (G.gate_matrix("CSWAP"), [9, 0, 11]),
             (G.gate_matrix("CCNOT"), [0, 1, 8]), (G.gate_matrix("SWAP"), [3, 6]), (G.gate_matrix("CNOT"), [3, 12]),
             (G.gate_matrix("PHASE", [1.3]), [2]), (ccp, [2, 5, 10])]  # bit 2 only inside: changed halves stored
    for U, qs in cases:
        say(f"  {dtype} {U.shape[0]}x{U.shape[0]} on {qs}")
        st.apply_matrix(U, G.nat_tuple(*qs))
        oracle.apply_matrix(ref, U, G.nat_tuple(*qs))
        check(st, ref, dtype)
    st.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_transpositions_with_every_bit_from_6_up(S, oracle, dtype):
    """FP64 single transpositions whose bits are all >= 6 run in the LDS
    exchange form (k_perm_xchg): CNOT both ways, SWAP, CCNOT, CSWAP.  Bit-exact
    against the oracle; FP32 takes the pair form."""
    n = 13
    rng = np.random.default_rng(67)
    say(f"exchange-form transpositions {dtype}")
    psi = rand_state(n, rng)
    st = S(n, dtype)
    st.write_amplitudes(psi)
    ref = psi.astype(np.complex64).astype(np.complex128) if dtype == "f32" else psi.copy()
    cases = [("CNOT", [7, 12]), ("CNOT", [12, 6]), ("SWAP", [6, 11]), ("CCNOT", [8, 10, 12]),
             ("CSWAP", [12, 6, 9]), ("SWAP", [9, 10])]
    for name, qs in cases:
        U = G.gate_matrix(name)
        st.apply_matrix(U, G.nat_tuple(*qs))
        oracle.apply_matrix(ref, U, G.nat_tuple(*qs))
        check(st, ref, dtype)
    st.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_general_permutations_by_cycles(S, oracle, dtype):
    """Permutation gates with several moved entries (3-cycles, a 4-cycle plus
    a transposition, random permutations of 3 qubits, a controlled 3-cycle)
    run cycle by cycle with one kept value, on random qubits and on fixed
    sets whose moved amplitudes come in long runs (all bits >= 5) or short
    ones (bits 0..2); a singular 0/1 matrix (two rows taking one column) goes
    to the dense kernel.  A permutation only moves amplitudes, so both dtypes
    are bit-exact against the oracle (FP32 starts from an FP32-rounded
    state)."""
    rng = np.random.default_rng(57)
    n = 13
    psi = rand_state(n, rng)
    if dtype == "f32":
        psi = psi.astype(np.complex64).astype(np.complex128)
    st = S(n, dtype)
    st.write_amplitudes(psi)
    ref = psi.copy()

    def pmat(perm):
        P = np.zeros((len(perm), len(perm)), dtype=np.complex128)
        for j, p in enumerate(perm):
            P[j, p] = 1
        return P

    mats = [pmat([1, 2, 0, 3]), pmat([3, 0, 1, 2]), pmat([1, 0, 3, 2, 5, 6, 7, 4])]
    mats += [pmat(list(rng.permutation(8))) for _ in range(4)]
    ctrl3 = np.eye(8, dtype=np.complex128)
    ctrl3[4:, 4:] = pmat([1, 2, 0, 3])
    mats.append(ctrl3)
    sing = np.array([[1, 0], [1, 0]], dtype=np.complex128)
    cases = [(U, None) for U in mats + [sing]]
    cases += [(mats[0], [7, 12]), (mats[2], [5, 11, 8]), (ctrl3, [9, 6, 12]),
              (mats[0], [1, 0]), (mats[2], [2, 0, 1]), (ctrl3, [0, 2, 1])]
    for U, fixed in cases:
        k = int(U.shape[0]).bit_length() - 1
        qs = fixed if fixed is not None else [int(x) for x in rng.choice(n, k, replace=False)]
        st.apply_matrix(U, qs)
        oracle.apply_matrix(ref, U, qs)
        got = st.amplitudes()
        if dtype == "f32":
            # the singular matrix is arithmetic (1*a + 0*b): FP32-exact here too
            got = np.asarray(got, dtype=np.complex128)
        np.testing.assert_array_equal(got, ref)
    st.close()
