"""Full-size GPU checks through size-independent properties: at the bench's
sizes (30 qubits FP64, 16 GiB; 32 qubits FP32, 32 GiB) the oracle is too slow,
so the tests use closed forms and invariants that a dropped term, a wrong
index or a transposed operand would break: a circuit followed by its inverse
returns |0...0> (plain, tiled and fused paths), norms stay 1, and QFT-30 of
|0> is the uniform state (amplitude 2^-15 everywhere)."""
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
