"""Full-size GPU checks through size-independent properties: at the bench's
sizes (30 qubits FP64, 16 GiB; 32 qubits FP32, 32 GiB) the oracle is too slow,
so the tests use closed forms and invariants that a dropped term, a wrong
index or a transposed operand would break: a circuit followed by its inverse
returns |0...0> (plain, tiled and fused paths), norms stay 1, QFT of |0> is
the uniform state (amplitude 2^-n/2 everywhere, read at offsets above 2^32 at
n = 34), and sampling a uniform state whose probabilities are exact
(Hadamard pairs, entries +-1/2) reproduces the closed form of its CDF,
C(i) = (i + 1) / 2^n, bit for bit under both tie rules."""
import sys

import numpy as np
import pytest

from qvm_amd import _native as N
from qvm_amd import circuits as CI
from qvm_amd import gates as G
from qvm_amd import planner as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def inverse(gates):
    return [(G.dagger(U), q) for U, q in reversed(gates)]


def mixed_circuit(n, seed):
    """Layered 1q/CNOT gates plus dense 2q/3q unitaries on far-apart qubits."""
    rng = np.random.default_rng(seed)
    gates = CI.layered(n, 3, seed=seed)
    for _ in range(12):
        k = int(rng.integers(2, 4))
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        Z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        gates.append((np.linalg.qr(Z)[0], qs))
    gates += CI.qft(n)[:80]
    return gates


def check_zero_state(st, tol):
    """|0> within tol, checked on the device: the norm, the first amplitude and
    the total probability elsewhere."""
    a0 = complex(st.read_amplitudes(0, 1)[0])
    total = st.total_probability()
    assert abs(abs(a0) - 1) < tol, a0
    assert abs(total - 1) < tol
    assert total - abs(a0) ** 2 < tol


@pytest.fixture(scope="module")
def N30():
    assert N.device_count() >= 1
    return 30


def test_plain_roundtrip_n30(N30):
    from qvm_amd.state import HipPureState
    n = N30
    gates = mixed_circuit(n, 1)
    say(f"full-size plain round trip n={n} gates={2 * len(gates)}")
    st = HipPureState(n, "f64")
    for U, q in gates + inverse(gates):
        st.apply_matrix(U, G.nat_tuple(*q))
    check_zero_state(st, 1e-12)
    st.close()


@pytest.mark.parametrize("mode", ["tiled_exact", "tiled_fma", "fused5"])
def test_planned_roundtrip_n30(N30, mode):
    from qvm_amd.state import HipPureState
    n = N30
    gates = mixed_circuit(n, 2)
    allg = gates + inverse(gates)
    if mode == "fused5":
        plan = P.plan_fused(allg, n, 5)
    else:
        plan = P.plan_tiled(allg, n, P.default_tile_m("f64"), 4, merge_1q=(mode == "tiled_fma"),
                            reorder=(mode == "tiled_fma"))
    say(f"full-size {mode} round trip n={n} items={len(plan)}")
    st = HipPureState(n, "f64", fma_tiles=(mode == "tiled_fma"))
    prog = P.build_program(st, plan, use_graph=True)
    prog.run()
    check_zero_state(st, 1e-11)
    prog.close()
    st.close()


def closed_form(u, n, rule):
    """Index the uniform CDF C(i) = (i + 1) / 2^n selects: rule "ge" is
    min{i : C(i) >= u} = ceil(u 2^n) - 1, "gt" min{i : C(i) > u} =
    floor(u 2^n); both clamped to [0, 2^n - 1] (u * 2^n is exact)."""
    x = np.asarray(u, dtype=np.float64) * float(1 << n)
    k = np.ceil(x) - 1 if rule == "ge" else np.floor(x)
    return np.clip(k, 0, (1 << n) - 1).astype(np.uint64)


def grid_draws(n, count, seed):
    """Draws placed exactly on CDF values k / 2^n (ties), plus 0 and 1."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 1 << n, count, dtype=np.int64)
    k[: min(4, count)] = [1, (1 << 32) + 7 if n > 33 else 2, (1 << n) - 1, 1 << (n - 1)][: min(4, count)]
    return np.concatenate([k.astype(np.float64) / float(1 << n), [0.0, 1.0]])


def hadamard_pairs(st, n):
    """H (x) H on qubit pairs (entries +-1/2, exact in any precision): every
    amplitude exactly 2^-n/2 (n even)."""
    HH = np.kron(G.gate_matrix("H"), G.gate_matrix("H"))
    HH = np.sign(HH.real) * 0.5 + 0j
    for q in range(0, n, 2):
        st.apply_matrix(HH, [q, q + 1])


def check_uniform_sampling(st, n, nshots, seed):
    u = np.concatenate([CI.uniforms(seed, nshots), grid_draws(n, 2000, seed)])
    for rule in ("ge", "gt"):
        got, near = st.sample_ex(u, rule, 1e-15)
        want = closed_form(u, n, rule)
        bad = np.nonzero(got != want)[0]
        say(f"  n={n} rule {rule}: {u.size} draws, {bad.size} mismatches, {int(near.sum())} near a boundary")
        assert bad.size == 0, (rule, bad[:5], u[bad[:5]], got[bad[:5]], want[bad[:5]])
        # every draw placed on a CDF value is flagged, and only such draws
        # (a uniform draw lands within 1e-15 of one of 2^n grid points with
        # probability 2^(n+1) 1e-15 each)
        assert near[nshots:-2].all()
    return u


def test_qft_uniform_n30(N30):
    from qvm_amd.state import HipPureState
    n = N30
    say(f"full-size QFT n={n}")
    st = HipPureState(n, "f64")
    for U, q in CI.qft(n):
        st.apply_matrix(U, G.nat_tuple(*q))
    # QFT|0> = 2^(-n/2) sum_x |x>: every amplitude real 2^-15 (sampled ranges)
    want = 2.0 ** (-n / 2)
    for start in (0, (1 << n) // 3, (1 << n) - 4096):
        a = st.read_amplitudes(start, 4096)
        assert np.max(np.abs(a - want)) < 1e-12
    assert abs(st.norm() - 1) < 1e-12
    # 10^6 shots against the closed form of the uniform CDF.  The QFT's
    # probabilities are 2^-30 to within rounding, so only draws within
    # rounding of a grid point k / 2^30 may resolve one index apart.
    u = CI.uniforms(3031, 1_000_000)
    got, near = st.sample_ex(u, "ge", 1e-15)
    want_i = closed_form(u, n, "ge")
    d = got.astype(np.int64) - want_i.astype(np.int64)
    bad = np.nonzero(d)[0]
    off = np.abs(u[bad] * float(1 << n) - np.round(u[bad] * float(1 << n)))
    say(f"  QFT-30 shots: {bad.size} mismatches vs closed form, {int(near.sum())} near a boundary, "
        f"max grid distance of a mismatch {off.max() if bad.size else 0:.3e} (index units)")
    assert np.abs(d).max() <= 1
    assert bad.size <= 16
    assert (off < 1e-3).all()
    st.close()


def test_uniform_sampling_exact_n30(N30):
    """The multi-segment prefix scan at the bench's size (2^22 chunks, 16
    segments per thread in the top-level scan) against the exact closed form."""
    from qvm_amd.state import HipPureState
    n = N30
    say(f"exact uniform sampling n={n} FP64")
    st = HipPureState(n, "f64")
    hadamard_pairs(st, n)
    assert st.read_amplitudes((1 << n) - 3, 1)[0] == 2.0 ** -15
    check_uniform_sampling(st, n, 200_000, 77)
    st.close()


def test_fp32_n34_uniform_sampling_and_indices_above_2_32():
    """Config 3's size (2^34 FP32 amplitudes, 128 GiB): amplitude reads,
    chunk prefix and draws at indices above 2^32, exact closed form."""
    from qvm_amd.state import HipPureState
    n = 34
    say(f"exact uniform sampling n={n} FP32")
    st = HipPureState(n, "f32")
    hadamard_pairs(st, n)
    for off in (0, (1 << 33) + 12345, (1 << 34) - 5):
        a = st.read_amplitudes(off, 4)
        assert np.all(a == 2.0 ** -17), (off, a)
    check_uniform_sampling(st, n, 100_000, 78)
    # measure-all collapses to |b> above 2^32 for a draw there
    u = ((1 << 33) + 12345.5) / float(1 << n)
    b = st.measure_all(u)
    assert b == (1 << 33) + 12345
    a = st.read_amplitudes(b - 1, 3)
    assert list(a) == [0, 1, 0]
    st.close()


def test_fp32_n34_qft_plain_and_tiled():
    """QFT-34 of |0> in FP32 on the plain and the tiled path: amplitudes read
    at offsets above 2^32 are 2^-17 to FP32 accuracy, and the paths agree."""
    from qvm_amd.state import HipPureState
    n = 34
    offs = (5, (1 << 32) + 3, (1 << 33) + 12345, (1 << 34) - 4096)
    got = {}
    for mode in ("plain", "tiled"):
        say(f"QFT-{n} FP32 {mode}")
        st = HipPureState(n, "f32", fma_tiles=(mode == "tiled"))
        gates = CI.qft(n)
        if mode == "plain":
            for U, q in gates:
                st.apply_matrix(U, G.nat_tuple(*q))
        else:
            P.execute_plan(st, P.plan_tiled(gates, n, P.default_tile_m("f32"), P.default_low("f32"),
                                            merge_1q=True))
        got[mode] = np.concatenate([st.read_amplitudes(o, 4096 if o + 4096 <= 1 << n else 4) for o in offs])
        assert abs(st.norm() - 1) < 1e-4
        st.close()
    want = 2.0 ** -17
    for mode, a in got.items():
        assert np.max(np.abs(a - want)) < want * 1e-4, mode
    assert np.max(np.abs(got["plain"] - got["tiled"])) < want * 1e-4


def test_fp32_n34_haar_cz_roundtrip():
    """Config 3's circuit family at n = 34 (Haar U3 + CZ matching, 3 layers)
    and its inverse, on the plain and tiled FP32 paths, back to |0>."""
    from qvm_amd.state import HipPureState
    n = 34
    gates = CI.haar_cz(n, 3, seed=21)
    allg = gates + inverse(gates)
    for mode in ("plain", "tiled"):
        say(f"Haar+CZ round trip n={n} FP32 {mode} gates={len(allg)}")
        st = HipPureState(n, "f32", fma_tiles=(mode == "tiled"))
        if mode == "plain":
            for U, q in allg:
                st.apply_matrix(U, G.nat_tuple(*q))
        else:
            P.execute_plan(st, P.plan_tiled(allg, n, P.default_tile_m("f32"), P.default_low("f32"),
                                            merge_1q=True))
        check_zero_state(st, 2e-5)
        st.close()


def test_fp32_fullsize_n32_roundtrip():
    from qvm_amd.state import HipPureState
    n = 32
    gates = CI.haar_cz(n, 3, seed=9)
    allg = gates + inverse(gates)
    plan = P.plan_tiled(allg, n, P.default_tile_m("f32"), P.default_low("f32"), merge_1q=True)
    say(f"full-size FP32 tiled round trip n={n} passes={len(plan)}")
    st = HipPureState(n, "f32", fma_tiles=True)
    P.execute_plan(st, plan)
    check_zero_state(st, 2e-5)
    st.close()
