"""GPU tests of the density-matrix machine (vec(rho) as a 2n-qubit engine
state) against dense numpy density matrices built from the same gate
matrices: rho -> U rho U^H, Kraus sums, measurement probabilities, collapse,
dephasing, measure-all."""
import numpy as np
import pytest

from qvm_amd import gates as G

pytestmark = pytest.mark.gpu


def full_op(U, quil_qubits, n):
    """Dense 2^n x 2^n operator of U on Quil-ordered qubits (bit j of the
    matrix index <-> nat-tuple qubit j)."""
    nat = G.nat_tuple(*quil_qubits)
    k = len(nat)
    F = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    for x in range(1 << n):
        col = sum(((x >> nat[j]) & 1) << j for j in range(k))
        for row in range(1 << k):
            y = x
            for j in range(k):
                y = (y & ~(1 << nat[j])) | (((row >> j) & 1) << nat[j])
            F[y, x] += U[row, col]
    return F


def test_unitary_evolution_is_psi_psi_dagger():
    from qvm_amd.density import make_density_qvm
    n = 4
    prog = "H 0\nCNOT 0 1\nRX(0.3) 2\nCPHASE(1.1) 2 3\nRY(0.7) 3\nCZ 1 3\nRZ(0.4) 0"
    q = make_density_qvm(n)
    q.load_program(prog).run()
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for line in prog.splitlines():
        name, qs = line.split(" ", 1)
        params = []
        if "(" in name:
            name, p = name.split("(")
            params = [float(p.rstrip(")"))]
        psi = full_op(G.gate_matrix(name, params), [int(x) for x in qs.split()], n) @ psi
    rho = np.outer(psi, psi.conj())
    assert np.max(np.abs(q.density_matrix() - rho)) < 1e-14
    np.testing.assert_allclose(q.measurement_probabilities(), np.abs(psi) ** 2, atol=1e-15)
    q.close()


def test_kraus_channel_and_dephasing():
    from qvm_amd.density import DensityQVM
    n = 3
    gamma = 0.3
    K0 = np.array([[1, 0], [0, np.sqrt(1 - gamma)]], dtype=np.complex128)
    K1 = np.array([[0, np.sqrt(gamma)], [0, 0]], dtype=np.complex128)
    q = DensityQVM(n)
    q.set_noisy_gate("X", [1], [K0 @ G.gate_matrix("X"), K1 @ G.gate_matrix("X")])
    q.load_program("H 0\nH 1\nCNOT 1 2\nX 1\nRY(0.4) 2\nMEASURE 0")
    q.run()
    rho = np.zeros((8, 8), dtype=np.complex128)
    rho[0, 0] = 1
    for name, qs, par in (("H", [0], []), ("H", [1], []), ("CNOT", [1, 2], [])):
        F = full_op(G.gate_matrix(name, par), qs, n)
        rho = F @ rho @ F.conj().T
    Ks = [full_op(K @ G.gate_matrix("X"), [1], n) for K in (K0, K1)]
    rho = sum(K @ rho @ K.conj().T for K in Ks)
    F = full_op(G.gate_matrix("RY", [0.4]), [2], n)
    rho = F @ rho @ F.conj().T
    # MEASURE 0 without a target drops the coherences of qubit 0
    idx = np.arange(8)
    keep = (((idx[:, None] ^ idx[None, :]) & 1) == 0)
    rho = np.where(keep, rho, 0)
    assert np.max(np.abs(q.density_matrix() - rho)) < 1e-14
    assert abs(np.trace(q.density_matrix()) - 1) < 1e-14
    assert q.draws_used == 0
    q.close()


def test_measure_collapse_and_measure_all():
    from qvm_amd.density import DensityQVM
    n = 3
    draws = iter([0.2, 0.9, 0.4, 0.7])
    q = DensityQVM(n, rng=lambda: next(draws))
    q.load_program("H 0\nRY(1.0) 1\nCNOT 0 2\nDECLARE ro BIT[1]\nMEASURE 0 ro[0]")
    q.run()
    psi = np.zeros(8, dtype=np.complex128)
    psi[0] = 1
    for name, qs, par in (("H", [0], []), ("RY", [1], [1.0]), ("CNOT", [0, 2], [])):
        psi = full_op(G.gate_matrix(name, par), qs, n) @ psi
    p1 = float(np.sum(np.abs(psi[1::2]) ** 2))
    bit = 1 if 0.2 <= p1 else 0
    assert int(q.memory_ref("ro", 0)) == bit
    proj = np.array([((i & 1) == bit) for i in range(8)], dtype=float)
    phi = proj * psi
    phi /= np.linalg.norm(phi)
    assert np.max(np.abs(q.density_matrix() - np.outer(phi, phi.conj()))) < 1e-14
    # measure-all: qubits 2, 1, 0 in turn (draws 0.9, 0.4, 0.7); the result is a
    # basis state consistent with the collapsed distribution
    _, bits = q.measure_all()
    b = sum(x << i for i, x in enumerate(bits))
    assert abs(phi[b]) > 0
    rho = q.density_matrix()
    assert abs(rho[b, b] - 1) < 1e-14 and abs(np.trace(rho) - 1) < 1e-14
    assert bits[0] == bit and bits[2] == bit  # collapsed qubit 0, and qubit 2 = qubit 0 (CNOT)
    assert q.draws_used <= 4
    q.close()
