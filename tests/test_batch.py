"""Batched QAOA runs (one captured graph for many instances) against the
oracle, instance by instance."""
import numpy as np
import pytest

from qvm_amd import circuits as CI
from qvm_amd import gates as G
from tests.sampling_check import check_draws

pytestmark = pytest.mark.gpu


def oracle_state(oracle, n, gates):
    psi = oracle.zero_state(n)
    for U, q in gates:
        oracle.apply_matrix(psi, U, G.nat_tuple(*q))
    return psi


def zz_expectation(psi, edges):
    idx = np.arange(psi.size)
    p = np.abs(psi) ** 2
    return sum(float(np.sum(p * (1 - 2 * (((idx >> a) ^ (idx >> b)) & 1)))) for a, b in edges)


@pytest.mark.parametrize("fma", [False, True])
def test_qaoa_batch_matches_oracle(oracle, fma):
    from qvm_amd.batch import qaoa_batch
    inst = CI.qaoa_instances(10, seed=4, nmin=10, nmax=16)
    shots = 512
    b, circuits, u = qaoa_batch(inst, "f64", shots=shots, seed=3, fma_tiles=fma)
    b.run()
    hc, samples = b.results()
    b.run()  # graph replay: identical results
    hc2, samples2 = b.results()
    assert np.array_equal(hc, hc2) and np.array_equal(samples, samples2)
    for i, (n, edges, gam, bet) in enumerate(inst):
        print(f"instance {i}: n={n} p={len(gam)} edges={len(edges)}", flush=True)
        ref = oracle_state(oracle, n, circuits[i])
        got = b.amplitudes(i)
        assert np.max(np.abs(got - ref)) < 1e-13
        assert abs(hc[i] - zz_expectation(ref, edges)) < 1e-12
        ui = u[i * shots:(i + 1) * shots]
        want = oracle.sample_multiple(ref, ui)
        check_draws(oracle.cdf(ref), ui, samples[i], want)
    b.close()


def test_qaoa_batch_large_instances_match_oracle(oracle):
    """Instances of 18, 20 and 22 qubits (the upper half of the bench's
    range, several tile passes each, FMA tiles at the autotuned width's
    range) against the oracle: amplitudes, <H_C> and 1024 shots each."""
    from qvm_amd.batch import qaoa_batch
    rng = CI.SplitMix64(17)
    inst = []
    for n in (18, 20, 22):
        edges = CI.random_regular3(n, rng)
        p = 2 + rng.below(3)
        inst.append((n, edges, [3.0 * rng.uniform() for _ in range(p)], [1.5 * rng.uniform() for _ in range(p)]))
    shots = 1024
    for m in (10, 12):
        b, circuits, u = qaoa_batch(inst, "f64", shots=shots, seed=31, fma_tiles=True, tile_m=m)
        b.run()
        hc, samples = b.results()
        for i, (n, edges, gam, _) in enumerate(inst):
            say(f"large batch instance n={n} p={len(gam)} m={m}")
            ref = oracle_state(oracle, n, circuits[i])
            assert np.max(np.abs(b.amplitudes(i) - ref)) < 1e-12
            assert abs(hc[i] - zz_expectation(ref, edges)) < 1e-11
            ui = u[i * shots:(i + 1) * shots]
            r = check_draws(oracle.cdf(ref), ui, samples[i], oracle.sample_multiple(ref, ui))
            say(f"  shots: {r}")
        b.close()


def test_qaoa_batch_f32_and_eager(oracle):
    from qvm_amd.batch import qaoa_batch
    inst = CI.qaoa_instances(6, seed=5, nmin=10, nmax=14)
    b64, circuits, _ = qaoa_batch(inst, "f64", shots=0, use_graph=False)
    b64.run()
    h64, _ = b64.results()
    b32, _, _ = qaoa_batch(inst, "f32", shots=0, fma_tiles=True)
    b32.run()
    h32, _ = b32.results()
    assert np.max(np.abs(h64 - h32)) < 1e-4
    for i, (n, edges, _, _) in enumerate(inst):
        ref = oracle_state(oracle, n, circuits[i])
        assert abs(h64[i] - zz_expectation(ref, edges)) < 1e-12
    b64.close()
    b32.close()


@pytest.mark.parametrize("gate_first", [False, True])
def test_batch_initial_state_each_run(oracle, gate_first):
    """Every run starts from |0...0>: through the first tile pass when every
    program begins with one (no reset sweep), through the reset sweep when
    one program begins with a standalone gate (a program of 3-qubit gates
    only, which the planner keeps out of tile passes)."""
    from qvm_amd.batch import Batch
    rng = np.random.default_rng(8)
    ns = [9, 12, 10]
    circuits = []
    for n in ns:
        gates = CI.layered(n, 2, seed=int(rng.integers(100)))
        circuits.append(gates)
    if gate_first:
        circuits[1] = [(np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0], qs)
                       for qs in ([0, 5, 11], [3, 1, 7], [11, 2, 9])]
    b = Batch(circuits, ns, "f64", use_graph=True)
    for run in range(3):
        b.run()
        b.sync()
        for i, n in enumerate(ns):
            ref = oracle_state(oracle, n, circuits[i])
            assert np.max(np.abs(b.amplitudes(i) - ref)) < 1e-13, (run, i)
    b.close()


@pytest.mark.parametrize("fma", [False, True])
def test_batch_passes_skip_zero_blocks(oracle, fma):
    """Runs start from |0...0>, so a pass's blocks whose base has a qubit set
    that no earlier pass or gate touched are zero and skipped: states of 14..18
    qubits in 10-qubit tiles (several passes each), one with standalone
    3-qubit gates between its passes, every run against the oracle."""
    from qvm_amd.batch import Batch
    rng = np.random.default_rng(21)
    ns = [16, 18, 14]
    circuits = []
    for n in ns:
        circuits.append(list(CI.layered(n, 3, seed=int(rng.integers(100)))))
    U3 = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0]
    circuits[2] = circuits[2][:20] + [(U3, [13, 2, 9])] + circuits[2][20:] + [(U3, [0, 12, 7])]
    b = Batch(circuits, ns, "f64", use_graph=True, fma_tiles=fma, tile_m=10)
    for run in range(2):
        b.run()
        b.sync()
        for i, n in enumerate(ns):
            say(f"zero-block skip fma={fma} run={run} instance={i} n={n}")
            ref = oracle_state(oracle, n, circuits[i])
            assert np.max(np.abs(b.amplitudes(i) - ref)) < 1e-13, (run, i)
    b.close()


def say(*a):
    print(*a, flush=True)


def test_tile_width_autotuner(oracle):
    """The autotuner times every candidate width on a sample and the batch it
    picks matches the oracle."""
    from qvm_amd import batch as B
    inst = CI.qaoa_instances(8, seed=6, nmin=10, nmax=14)
    circuits = [CI.qaoa_maxcut(n, e, g, b) for n, e, g, b in inst]
    best, timings = B.autotune_tile_m(circuits, [n for n, _, _, _ in inst], "f64", sample_every=2)
    assert set(timings) == set(B.tile_candidates("f64")) and best in timings
    assert all(t > 0 for t in timings.values())
    b, circuits, _ = B.qaoa_batch(inst, "f64", shots=0, fma_tiles=True, tile_m="auto")
    assert b.tile_m in timings and b.tile_timings is not None
    b.run()
    b.sync()
    for i, (n, _, _, _) in enumerate(inst):
        ref = oracle_state(oracle, n, circuits[i])
        assert np.max(np.abs(b.amplitudes(i) - ref)) < 1e-13
    b.close()
