"""HBM-resident pure state: the engine behind qvm's ``pure-state`` protocol.

``HipPureState`` plays the role of a new quantum STATE class in qvm
(src/state-representation.lisp:58-74 defines the CPU one; the protocol's
generics are at src/state-representation.lisp:25-47).  Amplitudes live in
device memory; ``amplitudes()`` / ``state_elements`` is the only device-to-host
copy, matching the reference's read-back points (src/qvm.lisp:63-82).

Matrix convention (all ``apply_*`` methods): row-major 2^k x 2^k complex
matrix whose index bit j selects ``qubits[j]``, i.e. ``qubits`` is in
nat-tuple order -- the reverse of the Quil argument order
(src/utilities.lisp:43-51).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _dtype_code(dtype) -> int:
    if dtype in (N.F64, "f64", "float64", "complex128", np.complex128):
        return N.F64
    if dtype in (N.F32, "f32", "float32", "complex64", np.complex64):
        return N.F32
    raise ValueError(f"unsupported dtype {dtype!r}; use 'f64' or 'f32'")


class HipPureState:
    """A 2^n-amplitude statevector on one GPU (|0...0> on creation)."""

    def __init__(self, num_qubits: int, dtype="f64", device: int = -1, memory_cap_bytes: int = 0,
                 fma_tiles: bool = False):
        """fma_tiles: tile passes contract complex products with fused
        multiply-adds (faster; no longer bit-identical to gate-by-gate)."""
        L = N.lib()
        self._h = C.c_void_p(None)
        cfg = N.Config(C.sizeof(N.Config), int(device), int(memory_cap_bytes),
                       N.FLAG_FMA_TILES if fma_tiles else 0)
        N.check(L.qvmhip_create(int(num_qubits), _dtype_code(dtype), C.byref(cfg), C.byref(self._h)))
        self.num_qubits = int(num_qubits)
        self.dtype = "f64" if _dtype_code(dtype) == N.F64 else "f32"
        self.fma_tiles = bool(fma_tiles)

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            N.lib().qvmhip_destroy(self._h)
            self._h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def length(self) -> int:
        return 1 << self.num_qubits

    def sync(self) -> None:
        N.check(N.lib().qvmhip_sync(self._h))

    # -- set-to-zero-state (src/wavefunction.lisp:79-91) ----------------
    def reset_zero(self) -> None:
        N.check(N.lib().qvmhip_reset_zero(self._h))

    # -- gate application -------------------------------------------------
    def _mat_call(self, fn, U, qubits):
        U = N.cmat(U)
        q = N.u32arr(qubits)
        k = len(q)
        if U.shape != (1 << k, 1 << k):
            raise ValueError(f"matrix shape {U.shape} does not match {k} qubits")
        N.check(fn(self._h, N.dptr(U.view(np.float64)), k, N.u32ptr(q)))

    def apply_matrix(self, U, qubits) -> None:
        """apply-matrix-operator (src/wavefunction.lisp:308-330)."""
        self._mat_call(N.lib().qvmhip_apply_matrix, U, qubits)

    def apply_matrix_dense(self, U, qubits) -> None:
        self._mat_call(N.lib().qvmhip_apply_matrix_dense, U, qubits)

    def apply_fused(self, U, qubits) -> None:
        self._mat_call(N.lib().qvmhip_apply_fused, U, qubits)

    def apply_diagonal(self, d, qubits) -> None:
        d = N.cmat(d).reshape(-1)
        q = N.u32arr(qubits)
        if d.size != 1 << len(q):
            raise ValueError("diagonal length does not match qubit count")
        N.check(N.lib().qvmhip_apply_diagonal(self._h, N.dptr(d.view(np.float64)), len(q), N.u32ptr(q)))

    def apply_permutation(self, perm, qubits) -> None:
        p = N.u32arr(perm)
        q = N.u32arr(qubits)
        N.check(N.lib().qvmhip_apply_permutation(self._h, N.u32ptr(p), len(q), N.u32ptr(q)))

    def apply_tile_pass(self, tile_qubits, ops) -> None:
        tq = N.u32arr(sorted(tile_qubits))
        arr = (N.TileOp * max(len(ops), 1))(*ops)
        N.check(N.lib().qvmhip_apply_tile_pass(self._h, len(tq), N.u32ptr(tq), len(ops), arr))

    # -- reductions -------------------------------------------------------
    def prob_excited(self, q: int) -> float:
        out = C.c_double()
        N.check(N.lib().qvmhip_prob_excited(self._h, int(q), C.byref(out)))
        return out.value

    def prob_ground(self, q: int) -> float:
        out = C.c_double()
        N.check(N.lib().qvmhip_prob_ground(self._h, int(q), C.byref(out)))
        return out.value

    def norm(self) -> float:
        out = C.c_double()
        N.check(N.lib().qvmhip_norm(self._h, C.byref(out)))
        return out.value

    def total_probability(self) -> float:
        """sum_i |psi_i|^2 (no square root)."""
        out = C.c_double()
        N.check(N.lib().qvmhip_total_probability(self._h, C.byref(out)))
        return out.value

    def normalize(self, length: float = 0.0) -> None:
        N.check(N.lib().qvmhip_normalize(self._h, float(length)))

    def probabilities(self, start: int = 0, count: int | None = None) -> np.ndarray:
        count = self.length - start if count is None else count
        out = np.empty(count, dtype=np.float64)
        N.check(N.lib().qvmhip_probabilities(self._h, start, count, N.dptr(out)))
        return out

    def expectation_zz(self, edges) -> float:
        e = N.u32arr(np.asarray(edges, dtype=np.uint32).reshape(-1))
        out = C.c_double()
        N.check(N.lib().qvmhip_expectation_zz(self._h, N.u32ptr(e), len(e) // 2, C.byref(out)))
        return out.value

    def inner_product(self, other: "HipPureState") -> complex:
        re, im = C.c_double(), C.c_double()
        N.check(N.lib().qvmhip_inner_product(self._h, other._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    # -- measurement ------------------------------------------------------
    def force_measure(self, q: int, bit: int, p1: float) -> None:
        """force-measurement (src/measurement.lisp:10-41)."""
        N.check(N.lib().qvmhip_force_measure(self._h, int(q), int(bit), float(p1)))

    def measure(self, q: int, u: float, compiled: bool = False):
        """Returns (bit, draws consumed); see src/measurement.lisp:87-105."""
        bit, used = C.c_int(), C.c_int()
        N.check(N.lib().qvmhip_measure(self._h, int(q), float(u), 1 if compiled else 0, C.byref(bit),
                                       C.byref(used)))
        return bit.value, used.value

    def measure_all(self, u: float) -> int:
        """measure-all-state (src/measurement.lisp:128-143): collapses."""
        out = C.c_uint64()
        N.check(N.lib().qvmhip_measure_all(self._h, float(u), C.byref(out)))
        return out.value

    def sample_measure_all(self, u) -> np.ndarray:
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
        out = np.zeros(u.size, dtype=np.uint64)
        N.check(N.lib().qvmhip_sample_measure_all(self._h, N.dptr(u), u.size, N.u64ptr(out)))
        return out

    def sample(self, u) -> np.ndarray:
        """sample-wavefunction-multiple-times (src/measurement.lisp:246-288)."""
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
        out = np.zeros(u.size, dtype=np.uint64)
        N.check(N.lib().qvmhip_sample(self._h, N.dptr(u), u.size, N.u64ptr(out)))
        return out

    def sample_ex(self, u, rule: str = "ge", eps: float = 1e-15):
        """Sampling with boundary reporting (qvmhip_sample_ex).

        rule "gt" is measure-all's C(b) > u (src/measurement.lisp:179-225),
        "ge" the multi-shot cdf >= u (:246-288).  Returns (indices, near)
        where near[i] is True when u[i] lies within eps of C(b) or C(b-1) as
        the engine's FP64 prefix formed them; near.sum() equals the count the
        engine reports.
        """
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
        out = np.zeros(u.size, dtype=np.uint64)
        near = np.zeros(u.size, dtype=np.uint8)
        count = C.c_uint64()
        r = N.RULE_GE if rule == "ge" else N.RULE_GT if rule == "gt" else None
        if r is None:
            raise ValueError("rule must be 'gt' or 'ge'")
        N.check(N.lib().qvmhip_sample_ex(self._h, N.dptr(u), u.size, r, float(eps), N.u64ptr(out),
                                         near.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(count)))
        flags = near.astype(bool)
        if int(flags.sum()) != count.value:
            raise RuntimeError("near-boundary count disagrees with the flags")
        return out, flags

    # -- amplitudes -------------------------------------------------------
    def read_amplitudes(self, start: int = 0, count: int | None = None) -> np.ndarray:
        count = self.length - start if count is None else count
        out = np.empty(count, dtype=np.complex128)
        N.check(N.lib().qvmhip_read_amplitudes(self._h, start, count, N.dptr(out.view(np.float64))))
        return out

    def wavefunction_octets(self, start: int = 0, count: int | None = None) -> bytes:
        """Amplitudes as the server's big-endian octet stream: re, im doubles,
        16 octets per amplitude (app/src/utilities.lisp:60-75)."""
        count = self.length - start if count is None else count
        out = np.empty(16 * count, dtype=np.uint8)
        N.check(N.lib().qvmhip_read_wavefunction_be(self._h, start, count, out.ctypes.data_as(C.c_void_p)))
        return out.tobytes()

    def probabilities_octets(self, start: int = 0, count: int | None = None) -> bytes:
        """|psi_i|^2 as big-endian doubles, 8 octets each (the probabilities
        reply, app/src/handle-request.lisp:154-175)."""
        count = self.length - start if count is None else count
        out = np.empty(8 * count, dtype=np.uint8)
        N.check(N.lib().qvmhip_read_probabilities_be(self._h, start, count, out.ctypes.data_as(C.c_void_p)))
        return out.tobytes()

    def write_amplitudes(self, amps, start: int = 0) -> None:
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128))
        N.check(N.lib().qvmhip_write_amplitudes(self._h, start, a.size, N.dptr(a.view(np.float64))))

    def amplitudes(self) -> np.ndarray:
        return self.read_amplitudes()

    def max_abs_diff(self, other: "HipPureState") -> float:
        out = C.c_double()
        N.check(N.lib().qvmhip_max_abs_diff(self._h, other._h, C.byref(out)))
        return out.value

    def tv_distance(self, other: "HipPureState") -> float:
        """Total-variation distance between the two outcome distributions."""
        out = C.c_double()
        N.check(N.lib().qvmhip_tv_distance(self._h, other._h, C.byref(out)))
        return out.value

    # -- live kernel timing ----------------------------------------------
    def profile(self, on: bool = True) -> None:
        N.check(N.lib().qvmhip_profile_enable(self._h, 1 if on else 0))

    def profile_summary(self) -> dict:
        """{family: (launches, total ms)} of the launches recorded since
        profile(True)."""
        out = {}
        for k, name in N.K_NAMES.items():
            c, t = C.c_uint64(), C.c_double()
            N.check(N.lib().qvmhip_profile_summary(self._h, k, C.byref(c), C.byref(t)))
            if c.value:
                out[name] = (c.value, t.value)
        return out

    def copy_from(self, other: "HipPureState") -> None:
        """copy-wavefunction (src/wavefunction.lisp:93-109)."""
        N.check(N.lib().qvmhip_copy(self._h, other._h))

    def accumulate(self, other: "HipPureState") -> None:
        """self += other, elementwise."""
        N.check(N.lib().qvmhip_accumulate(self._h, other._h))

    def density_diagonal(self, nq: int) -> np.ndarray:
        """Re rho[i][i] of the density matrix this 2*nq-qubit state holds
        as vec(rho), row-major."""
        out = np.empty(1 << nq, dtype=np.float64)
        N.check(N.lib().qvmhip_density_diagonal(self._h, int(nq), N.dptr(out)))
        return out

    def swap_contents(self, other: "HipPureState") -> None:
        """Exchange the device buffers of two states of one size and dtype
        (the rotatef of the state and trial amplitude vectors,
        src/apply-gate.lisp:37).  Programs built on either state keep the
        buffer they were built on."""
        if other.num_qubits != self.num_qubits or other.dtype != self.dtype:
            raise ValueError("states differ in size or dtype")
        self._h, other._h = other._h, self._h
        self.fma_tiles, other.fma_tiles = other.fma_tiles, self.fma_tiles


class Program:
    """A compiled gate list replayed on one state (optionally a hipGraph)."""

    def __init__(self, state: HipPureState):
        self.state = state
        self._h = C.c_void_p(None)
        N.check(N.lib().qvmhip_program_create(state.handle, C.byref(self._h)))

    def add_matrix(self, U, qubits) -> None:
        U = N.cmat(U)
        q = N.u32arr(qubits)
        N.check(N.lib().qvmhip_program_add_matrix(self._h, N.dptr(U.view(np.float64)), len(q), N.u32ptr(q)))

    def add_fused(self, U, qubits) -> None:
        U = N.cmat(U)
        q = N.u32arr(qubits)
        N.check(N.lib().qvmhip_program_add_fused(self._h, N.dptr(U.view(np.float64)), len(q), N.u32ptr(q)))

    def add_tile_pass(self, tile_qubits, ops) -> None:
        tq = N.u32arr(sorted(tile_qubits))
        arr = (N.TileOp * max(len(ops), 1))(*ops)
        N.check(N.lib().qvmhip_program_add_tile_pass(self._h, len(tq), N.u32ptr(tq), len(ops), arr))

    def finalize(self, use_graph: bool = True) -> None:
        N.check(N.lib().qvmhip_program_finalize(self._h, 1 if use_graph else 0))

    def run(self) -> None:
        N.check(N.lib().qvmhip_program_run(self._h))

    def time(self, reps: int = 1) -> float:
        ms = C.c_double()
        N.check(N.lib().qvmhip_time_program(self._h, int(reps), C.byref(ms)))
        return ms.value

    def close(self) -> None:
        if self._h and self._h.value:
            N.lib().qvmhip_program_destroy(self._h)
            self._h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
