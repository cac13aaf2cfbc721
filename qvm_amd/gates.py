"""Standard Quil gate matrices and gate modifiers.

The matrices are the formulas of the reference's gate library
(quil/stdgates.quil:1-206; e.g. RX :31, RZ :39, CNOT :46, CPHASE :96, CZ :102,
SWAP :110, RZZ :148).  They are restated here because cl-quil, which builds
them in the reference (``quil:gate-matrix``, called from
src/apply-gate.lisp:112-160), is not vendored.  Matrices use the Quil
convention: for a gate applied to qubits (a, b, ...), the FIRST argument is the
most significant bit of the matrix index; ``nat_tuple`` reverses the argument
list into the engine's order, where index bit j selects ``qubits[j]``
(src/utilities.lisp:43-51).
"""
from __future__ import annotations

import math

import numpy as np

SQRT1_2 = 1.0 / math.sqrt(2.0)


def cis(x: float) -> complex:
    return complex(math.cos(x), math.sin(x))


def _m(rows) -> np.ndarray:
    return np.array(rows, dtype=np.complex128)


def _rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _m([[c, -1j * s], [-1j * s, c]])


def _ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _m([[c, -s], [s, c]])


def _rz(t):
    return _m([[cis(-t / 2), 0], [0, cis(t / 2)]])


def _phase(a):
    return _m([[1, 0], [0, cis(a)]])


def _cphase(a):
    return np.diag([1, 1, 1, cis(a)]).astype(np.complex128)


def _cphase00(a):
    return np.diag([cis(a), 1, 1, 1]).astype(np.complex128)


def _cphase01(a):
    return np.diag([1, cis(a), 1, 1]).astype(np.complex128)


def _cphase10(a):
    return np.diag([1, 1, cis(a), 1]).astype(np.complex128)


def _pswap(t):
    return _m([[1, 0, 0, 0], [0, 0, cis(t), 0], [0, cis(t), 0, 0], [0, 0, 0, 1]])


def _rzz(p):
    return np.diag([cis(-p / 2), cis(p / 2), cis(p / 2), cis(-p / 2)]).astype(np.complex128)


def _rxx(p):
    c, s = math.cos(p / 2), -1j * math.sin(p / 2)
    return _m([[c, 0, 0, s], [0, c, s, 0], [0, s, c, 0], [s, 0, 0, c]])


def _ryy(p):
    c, s = math.cos(p / 2), 1j * math.sin(p / 2)
    return _m([[c, 0, 0, s], [0, c, -s, 0], [0, -s, c, 0], [s, 0, 0, c]])


def _xy(t):
    c, s = math.cos(t / 2), 1j * math.sin(t / 2)
    return _m([[1, 0, 0, 0], [0, c, s, 0], [0, s, c, 0], [0, 0, 0, 1]])


def _fsim(t, p):
    c, s = math.cos(t / 2), 1j * math.sin(t / 2)
    return _m([[1, 0, 0, 0], [0, c, s, 0], [0, s, c, 0], [0, 0, 0, cis(p)]])


def _u3(theta, phi, lam):
    """U3(theta, phi, lambda) in its standard closed form."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return _m([[c, -cis(lam) * s], [cis(phi) * s, cis(phi + lam) * c]])


_FIXED = {
    "I": _m([[1, 0], [0, 1]]),
    "X": _m([[0, 1], [1, 0]]),
    "Y": _m([[0, -1j], [1j, 0]]),
    "Z": _m([[1, 0], [0, -1]]),
    "H": _m([[SQRT1_2, SQRT1_2], [SQRT1_2, -SQRT1_2]]),
    "S": _m([[1, 0], [0, 1j]]),
    "T": _m([[1, 0], [0, cis(math.pi / 4)]]),
    "CNOT": _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]),
    "CZ": np.diag([1, 1, 1, -1]).astype(np.complex128),
    "SWAP": _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]),
    "ISWAP": _m([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
    "SQISWAP": _m([[1, 0, 0, 0], [0, SQRT1_2, 1j * SQRT1_2, 0], [0, 1j * SQRT1_2, SQRT1_2, 0], [0, 0, 0, 1]]),
}
_ccnot = np.eye(8, dtype=np.complex128)
_ccnot[[6, 7]] = _ccnot[[7, 6]]
_FIXED["CCNOT"] = _ccnot
_cswap = np.eye(8, dtype=np.complex128)
_cswap[[5, 6]] = _cswap[[6, 5]]
_FIXED["CSWAP"] = _cswap

_PARAM = {
    "RX": (1, _rx), "RY": (1, _ry), "RZ": (1, _rz), "PHASE": (1, _phase),
    "CPHASE": (1, _cphase), "CPHASE00": (1, _cphase00), "CPHASE01": (1, _cphase01),
    "CPHASE10": (1, _cphase10), "PSWAP": (1, _pswap), "RZZ": (1, _rzz), "RXX": (1, _rxx),
    "RYY": (1, _ryy), "XY": (1, _xy), "PISWAP": (1, _xy), "FSIM": (2, _fsim), "U3": (3, _u3),
}

STANDARD_GATES = sorted(set(_FIXED) | set(_PARAM))


def is_standard(name: str) -> bool:
    return name in _FIXED or name in _PARAM


def gate_matrix(name: str, params=()) -> np.ndarray:
    """The 2^k x 2^k matrix of a standard gate, Quil convention."""
    if name in _FIXED:
        if params:
            raise ValueError(f"gate {name} takes no parameters")
        return _FIXED[name].copy()
    if name in _PARAM:
        npar, fn = _PARAM[name]
        if len(params) != npar:
            raise ValueError(f"gate {name} takes {npar} parameter(s), got {len(params)}")
        return fn(*[float(np.real(p)) for p in params])
    raise KeyError(f"unknown gate {name}")


def num_qubits_of(U: np.ndarray) -> int:
    k = int(U.shape[0]).bit_length() - 1
    if U.shape != (1 << k, 1 << k):
        raise ValueError("gate matrix must be square with power-of-two size")
    return k


def dagger(U: np.ndarray) -> np.ndarray:
    """DAGGER modifier: conjugate transpose."""
    return np.ascontiguousarray(U.conj().T)


def controlled(U: np.ndarray) -> np.ndarray:
    """CONTROLLED modifier: the control is the new first (most significant) qubit."""
    d = U.shape[0]
    out = np.eye(2 * d, dtype=np.complex128)
    out[d:, d:] = U
    return out


def forked(U0: np.ndarray, U1: np.ndarray) -> np.ndarray:
    """FORKED modifier: first qubit selects U0 (|0>) or U1 (|1>)."""
    d = U0.shape[0]
    out = np.zeros((2 * d, 2 * d), dtype=np.complex128)
    out[:d, :d] = U0
    out[d:, d:] = U1
    return out


def nat_tuple(*qubits: int) -> tuple:
    """Reverse a Quil argument list into engine order (src/utilities.lisp:43-51)."""
    return tuple(reversed(qubits))


def changed_amplitudes(U: np.ndarray, n: int) -> int:
    """Amplitudes gate U (k qubits) on an n-qubit state can modify:
    2^(n-k) * #{j : row j of U != e_j}.  Dense 1q gates: 2^n, CNOT/SWAP:
    2^(n-1), CZ/CPHASE: 2^(n-2)."""
    k = num_qubits_of(U)
    eye = np.eye(U.shape[0], dtype=np.complex128)
    moved = int(np.sum(np.any(U != eye, axis=1)))
    return moved << (n - k)


def is_unitary(U: np.ndarray, tol: float = 1e-10) -> bool:
    return np.allclose(U.conj().T @ U, np.eye(U.shape[0]), atol=tol)
