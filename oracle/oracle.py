"""ctypes wrapper of the CPU oracle (oracle/qvm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker of the HIP engine.
The product package ``qvm_amd`` never imports this module.

Every function mirrors one function of quil-lang/qvm 1.18.0 (citations are in
qvm_oracle.c).  Amplitudes are numpy ``complex128`` vectors (interleaved
re/im, the in-memory layout of SBCL's ``(complex double-float)`` arrays).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libqvm_oracle.so")
_lib = None

_u64 = C.c_uint64
_u32 = C.c_uint32
_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_set_parallelization": (None, [C.c_int, C.c_int]),
            "orc_count_logical_cores": (C.c_int, []),
            "orc_inject_bit": (_u64, [_u64, C.c_uint]),
            "orc_eject_bit": (_u64, [_u64, C.c_uint]),
            "orc_index_to_address": (_u64, [_u64, C.c_uint, C.c_uint]),
            "orc_nat_tuple": (None, [_u32p, _u32, _u32p]),
            "orc_subdivide": (C.c_int, [_u64, C.c_int, _u64p, _u64p]),
            "orc_bring_to_zero_state": (None, [_dp, _u64]),
            "orc_apply_matrix": (C.c_int, [_dp, _u32, _dp, _u32, _u32p]),
            "orc_apply_compiled": (C.c_int, [_dp, _u32, _dp, _u32, _u32p]),
            "orc_apply_avx": (C.c_int, [_dp, _u32, _dp, _u32, _u32p]),
            "orc_permutation_to_transpositions": (C.c_int, [_u32p, _u32, _u32p, _u32p]),
            "orc_apply_permutation": (C.c_int, [_dp, _u32, _u32p, _u32, _u32p]),
            "orc_excited_probability": (C.c_double, [_dp, _u32, _u32]),
            "orc_ground_probability": (C.c_double, [_dp, _u32, _u32]),
            "orc_norm": (C.c_double, [_dp, _u64]),
            "orc_normalize": (None, [_dp, _u64, C.c_double]),
            "orc_force_measurement": (None, [_dp, _u32, _u32, C.c_int, C.c_double]),
            "orc_measure": (C.c_int, [_dp, _u32, _u32, C.c_double, C.c_int, C.POINTER(C.c_int)]),
            "orc_sample_bisect": (_u64, [_dp, _u64, C.c_double]),
            "orc_measure_all": (_u64, [_dp, _u32, C.c_double]),
            "orc_cdf": (None, [_dp, _u64, _dp]),
            "orc_cdf_search": (_u64, [_dp, _u64, C.c_double]),
            "orc_sample_multiple": (None, [_dp, _u64, _dp, _u64, _u64p, _dp]),
            "orc_sample_as_distribution": (None, [_dp, _u64, _dp, _u64, _u64p]),
            "orc_copy": (None, [_dp, _dp, _u64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _cvec(psi: np.ndarray) -> np.ndarray:
    assert psi.dtype == np.complex128 and psi.flags["C_CONTIGUOUS"], "oracle needs contiguous complex128"
    return psi


def _qubits(qs) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(qs, dtype=np.uint32))


def nq_of(psi: np.ndarray) -> int:
    n = int(psi.size).bit_length() - 1
    assert 1 << n == psi.size
    return n


# --- configuration --------------------------------------------------------
def set_parallelization(log2_threshold: int = 19, workers: int = 0) -> None:
    lib().orc_set_parallelization(int(log2_threshold), int(workers))


def count_logical_cores() -> int:
    return int(lib().orc_count_logical_cores())


# --- index math -----------------------------------------------------------
def inject_bit(x: int, n: int) -> int:
    return int(lib().orc_inject_bit(x, n))


def eject_bit(x: int, n: int) -> int:
    return int(lib().orc_eject_bit(x, n))


def index_to_address(index: int, qubit: int, state: int) -> int:
    return int(lib().orc_index_to_address(index, qubit, state))


def nat_tuple(*args: int) -> list:
    a = _qubits(args)
    out = np.zeros_like(a)
    lib().orc_nat_tuple(a.ctypes.data_as(_u32p), len(a), out.ctypes.data_as(_u32p))
    return [int(x) for x in out]


def subdivide(total: int, num_ranges: int) -> list:
    s = np.zeros(max(num_ranges, 1), dtype=np.uint64)
    e = np.zeros_like(s)
    n = lib().orc_subdivide(total, num_ranges, s.ctypes.data_as(_u64p), e.ctypes.data_as(_u64p))
    return [(int(s[i]), int(e[i])) for i in range(n)]


def permutation_to_transpositions(perm) -> list:
    p = _qubits(perm)
    left = np.zeros(64, dtype=np.uint32)
    right = np.zeros(64, dtype=np.uint32)
    n = lib().orc_permutation_to_transpositions(
        p.ctypes.data_as(_u32p), len(p), left.ctypes.data_as(_u32p), right.ctypes.data_as(_u32p))
    return [(int(left[i]), int(right[i])) for i in range(n)]


# --- state ----------------------------------------------------------------
def zero_state(nq: int) -> np.ndarray:
    psi = np.empty(1 << nq, dtype=np.complex128)
    lib().orc_bring_to_zero_state(_dptr(psi.view(np.float64)), psi.size)
    return psi


def _apply(fn, psi, U, qubits):
    psi = _cvec(psi)
    U = np.ascontiguousarray(np.asarray(U, dtype=np.complex128))
    q = _qubits(qubits)
    k = len(q)
    assert U.shape == (1 << k, 1 << k)
    rc = fn(_dptr(psi.view(np.float64)), nq_of(psi), _dptr(U.view(np.float64)), k,
            q.ctypes.data_as(_u32p))
    if rc != 0:
        raise ValueError("oracle rejected gate arguments")
    return psi


def apply_matrix(psi, U, qubits):
    """Serial/generic kernel semantics; ``qubits`` in nat-tuple order."""
    return _apply(lib().orc_apply_matrix, psi, U, qubits)


def apply_compiled(psi, U, qubits):
    return _apply(lib().orc_apply_compiled, psi, U, qubits)


def apply_avx(psi, U, qubits):
    return _apply(lib().orc_apply_avx, psi, U, qubits)


def apply_permutation(psi, perm, qubits):
    psi = _cvec(psi)
    p = _qubits(perm)
    q = _qubits(qubits)
    rc = lib().orc_apply_permutation(_dptr(psi.view(np.float64)), nq_of(psi), p.ctypes.data_as(_u32p),
                                     len(q), q.ctypes.data_as(_u32p))
    if rc != 0:
        raise ValueError("oracle rejected permutation arguments")
    return psi


# --- reductions / measurement -------------------------------------------
def excited_probability(psi, q: int) -> float:
    return float(lib().orc_excited_probability(_dptr(_cvec(psi).view(np.float64)), nq_of(psi), q))


def ground_probability(psi, q: int) -> float:
    return float(lib().orc_ground_probability(_dptr(_cvec(psi).view(np.float64)), nq_of(psi), q))


def norm(psi) -> float:
    return float(lib().orc_norm(_dptr(_cvec(psi).view(np.float64)), psi.size))


def normalize(psi, length: float = 0.0):
    lib().orc_normalize(_dptr(_cvec(psi).view(np.float64)), psi.size, float(length))
    return psi


def force_measurement(psi, q: int, bit: int, p1: float):
    lib().orc_force_measurement(_dptr(_cvec(psi).view(np.float64)), nq_of(psi), q, bit, p1)
    return psi


def measure(psi, q: int, u: float, compiled: bool = False):
    consumed = C.c_int(0)
    bit = lib().orc_measure(_dptr(_cvec(psi).view(np.float64)), nq_of(psi), q, float(u),
                            1 if compiled else 0, C.byref(consumed))
    return int(bit), int(consumed.value)


def sample_bisect(psi, p: float) -> int:
    return int(lib().orc_sample_bisect(_dptr(_cvec(psi).view(np.float64)), psi.size, float(p)))


def measure_all(psi, p: float) -> int:
    return int(lib().orc_measure_all(_dptr(_cvec(psi).view(np.float64)), nq_of(psi), float(p)))


def cdf(psi) -> np.ndarray:
    out = np.empty(psi.size, dtype=np.float64)
    lib().orc_cdf(_dptr(_cvec(psi).view(np.float64)), psi.size, _dptr(out))
    return out


def cdf_search(c: np.ndarray, p: float) -> int:
    return int(lib().orc_cdf_search(_dptr(c), c.size, float(p)))


def sample_multiple(psi, u) -> np.ndarray:
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    out = np.zeros(u.size, dtype=np.uint64)
    scratch = np.empty(psi.size, dtype=np.float64)
    lib().orc_sample_multiple(_dptr(_cvec(psi).view(np.float64)), psi.size, _dptr(u), u.size,
                              out.ctypes.data_as(_u64p), _dptr(scratch))
    return out


def sample_as_distribution(psi, ps) -> np.ndarray:
    ps = np.ascontiguousarray(np.asarray(ps, dtype=np.float64))
    out = np.zeros(ps.size, dtype=np.uint64)
    lib().orc_sample_as_distribution(_dptr(_cvec(psi).view(np.float64)), psi.size, _dptr(ps), ps.size,
                                     out.ctypes.data_as(_u64p))
    return out
