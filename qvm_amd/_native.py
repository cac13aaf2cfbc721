"""ctypes binding of the qvmhip C ABI (include/qvmhip.h).

This module is the Python twin of the Lisp CFFI binding a qvm maintainer
would add: one foreign function per entry point, same argument order.  It
loads the in-tree ``qvm_amd/_lib/libqvmhip.so`` and raises if it is missing
or was built from other sources than the tree holds (the build writes a hash
of its sources next to the library); there is no CPU fallback anywhere in
the package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libqvmhip.so")

F64 = 0
F32 = 1
RULE_GT = 0  # sampling rules of qvmhip_sample_ex
RULE_GE = 1
FLAG_FMA_TILES = 1  # qvmhip_config.flags

OK = 0
EINVAL = -1
ENOMEM = -2
EDEVICE = -3
ERANGE = -4
EUNSUPPORTED = -5

TILE_DENSE1 = 1
TILE_DENSE2 = 2
TILE_DIAG1 = 3
TILE_DIAG2 = 4
TILE_CPERM = 5
TILE_SWAP = 6
TILE_CPHASE = 7
TILE_GDIAG = 8
# slot-space kinds emitted by the tile scheduler (qvmhip_tile_schedule)
TILE_XREG = 16
TILE_RSWAP = 17
TILE_REMAP = 18
TILE_RDIAG = 19
TILE_D1REAL = 20
TILE_D1RX = 21
TILE_CUT_PHASE = 22
TILE_DATA = 23


class QVMHipError(RuntimeError):
    """An error status returned by the engine; ``code`` is the status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"qvmhip error {code}: {message}")
        self.code = code


class TileOp(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("q0", C.c_uint32), ("q1", C.c_uint32),
                ("ctrl_mask", C.c_uint32), ("m", C.c_double * 32)]


class Config(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("device", C.c_int32),
                ("memory_cap_bytes", C.c_uint64), ("flags", C.c_uint32)]


_u32 = C.c_uint32
_u64 = C.c_uint64
_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "qvmhip_last_error": [],
    "qvmhip_version": [],
    "qvmhip_device_count": [_ip],
    "qvmhip_state_bytes": [_u32, C.c_int, _u64p],
    "qvmhip_create": [_u32, C.c_int, C.POINTER(Config), C.POINTER(_vp)],
    "qvmhip_destroy": [_vp],
    "qvmhip_num_qubits": [_vp, _u32p],
    "qvmhip_dtype": [_vp, _ip],
    "qvmhip_reset_zero": [_vp],
    "qvmhip_sync": [_vp],
    "qvmhip_apply_matrix": [_vp, _dp, _u32, _u32p],
    "qvmhip_apply_matrix_dense": [_vp, _dp, _u32, _u32p],
    "qvmhip_apply_diagonal": [_vp, _dp, _u32, _u32p],
    "qvmhip_apply_permutation": [_vp, _u32p, _u32, _u32p],
    "qvmhip_apply_fused": [_vp, _dp, _u32, _u32p],
    "qvmhip_apply_tile_pass": [_vp, _u32, _u32p, _u32, C.POINTER(TileOp)],
    "qvmhip_tile_schedule": [C.c_int, _u32, _u32, _u32, C.POINTER(TileOp), C.POINTER(TileOp), _u32, _u32p],
    "qvmhip_prob_excited": [_vp, _u32, _dp],
    "qvmhip_prob_ground": [_vp, _u32, _dp],
    "qvmhip_norm": [_vp, _dp],
    "qvmhip_total_probability": [_vp, _dp],
    "qvmhip_normalize": [_vp, C.c_double],
    "qvmhip_probabilities": [_vp, _u64, _u64, _dp],
    "qvmhip_expectation_zz": [_vp, _u32p, _u32, _dp],
    "qvmhip_inner_product": [_vp, _vp, _dp, _dp],
    "qvmhip_force_measure": [_vp, _u32, C.c_int, C.c_double],
    "qvmhip_measure": [_vp, _u32, C.c_double, C.c_int, _ip, _ip],
    "qvmhip_measure_all": [_vp, C.c_double, _u64p],
    "qvmhip_sample_measure_all": [_vp, _dp, _u64, _u64p],
    "qvmhip_sample": [_vp, _dp, _u64, _u64p],
    "qvmhip_sample_ex": [_vp, _dp, _u64, C.c_int, C.c_double, _u64p, C.POINTER(C.c_uint8), _u64p],
    "qvmhip_sample_noalloc": [C.c_void_p, _u32, C.c_int, C.c_void_p, _u64, C.c_int, C.c_double, C.c_void_p,
                              C.c_void_p, C.c_void_p],
    "qvmhip_read_amplitudes": [_vp, _u64, _u64, _dp],
    "qvmhip_accumulate": [_vp, _vp],
    "qvmhip_density_diagonal": [_vp, _u32, _dp],
    "qvmhip_read_wavefunction_be": [_vp, _u64, _u64, C.c_void_p],
    "qvmhip_read_probabilities_be": [_vp, _u64, _u64, C.c_void_p],
    "qvmhip_write_amplitudes": [_vp, _u64, _u64, _dp],
    "qvmhip_copy": [_vp, _vp],
    "qvmhip_program_create": [_vp, C.POINTER(_vp)],
    "qvmhip_program_add_matrix": [_vp, _dp, _u32, _u32p],
    "qvmhip_program_add_fused": [_vp, _dp, _u32, _u32p],
    "qvmhip_program_add_tile_pass": [_vp, _u32, _u32p, _u32, C.POINTER(TileOp)],
    "qvmhip_program_finalize": [_vp, C.c_int],
    "qvmhip_program_run": [_vp],
    "qvmhip_program_destroy": [_vp],
    "qvmhip_time_program": [_vp, C.c_int, _dp],
    "qvmhip_max_abs_diff": [_vp, _vp, _dp],
    "qvmhip_tv_distance": [_vp, _vp, _dp],
    "qvmhip_profile_enable": [_vp, C.c_int],
    "qvmhip_profile_summary": [_vp, C.c_int, _u64p, _dp],
    "qvmhip_batch_create": [_u32, _u32p, C.c_int, C.POINTER(_vp)],
    "qvmhip_batch_destroy": [_vp],
    "qvmhip_batch_set_flags": [_vp, _u32],
    "qvmhip_batch_add_tile_pass": [_vp, _u32, _u32, _u32p, _u32, C.POINTER(TileOp)],
    "qvmhip_batch_add_matrix": [_vp, _u32, _dp, _u32, _u32p],
    "qvmhip_batch_set_zz": [_vp, _u32, _u32p, _u32],
    "qvmhip_batch_set_shots": [_vp, _u64, _dp],
    "qvmhip_batch_finalize": [_vp, C.c_int],
    "qvmhip_batch_run": [_vp],
    "qvmhip_batch_sync": [_vp],
    "qvmhip_batch_results": [_vp, _dp, _u64p],
    "qvmhip_batch_read_amplitudes": [_vp, _u32, _u64, _u64, _dp],
    "qvmhip_mlp_create": [C.c_int, C.c_int, _fp, C.POINTER(_vp)],
    "qvmhip_mlp_destroy": [_vp],
    "qvmhip_mlp_num_params": [_vp, _ip],
    "qvmhip_mlp_get_params": [_vp, _fp],
    "qvmhip_mlp_set_params": [_vp, _fp],
    "qvmhip_mlp_predict": [_vp, _fp, C.c_int, _fp],
    "qvmhip_mlp_gradient": [_vp, _fp, _fp, C.c_int, _fp, _fp],
    "qvmhip_mlp_train": [_vp, _fp, _fp, C.c_int, C.c_int, C.c_float, _fp],
}

# kernel families of qvmhip_profile_summary
K_DENSE1, K_DENSE2, K_DENSE3, K_DENSE_BIG, K_DIAG_FULL, K_DIAG_SPARSE, K_PERM, K_TILE, K_FUSED = range(1, 10)
K_NAMES = {K_DENSE1: "dense1", K_DENSE2: "dense2", K_DENSE3: "dense3", K_DENSE_BIG: "dense_big",
           K_DIAG_FULL: "diag_full", K_DIAG_SPARSE: "diag_sparse", K_PERM: "perm", K_TILE: "tile",
           K_FUSED: "fused"}
_RESTYPE = {"qvmhip_last_error": C.c_char_p, "qvmhip_version": C.c_char_p}

_lib = None


def lib():
    """Load the engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"qvmhip library not built: {LIB_PATH} is missing (run __graft_entry__.build() "
                "or python qvm_amd/build.py)")
        from . import build as B
        stamp = open(B.STAMP).read().strip() if os.path.exists(B.STAMP) else ""
        if stamp != B.source_hash():
            raise ImportError(
                f"qvmhip library {LIB_PATH} was not built from the sources in this tree (its build stamp "
                "differs from their hash): rebuild with __graft_entry__.build() or python qvm_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def last_error() -> str:
    return lib().qvmhip_last_error().decode()


def check(rc: int) -> None:
    if rc != OK:
        raise QVMHipError(rc, last_error())


def device_count() -> int:
    n = C.c_int(0)
    check(lib().qvmhip_device_count(C.byref(n)))
    return n.value


def dptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def u32arr(xs) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(xs, dtype=np.uint32))


def u32ptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def u64ptr(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)


def cmat(U) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(U, dtype=np.complex128))


def tile_schedule(dtype: str, m: int, ops, fma: bool = False) -> list:
    """The slot-space program the engine runs for a tile pass (see
    qvmhip_tile_schedule)."""
    arr = (TileOp * max(1, len(ops)))(*ops)
    cap = 4 * len(ops) + 8
    out = (TileOp * cap)()
    n = C.c_uint32()
    check(lib().qvmhip_tile_schedule(F64 if dtype == "f64" else F32, FLAG_FMA_TILES if fma else 0, m, len(ops),
                                     arr, out, cap, C.byref(n)))
    return [out[i] for i in range(n.value)]
