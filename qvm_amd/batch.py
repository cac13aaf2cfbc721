"""Many small circuits in one captured launch sequence (QAOA parameter sweeps).

The reference runs a QAOA sweep one QVM at a time (examples/qaoa.lisp:7-75,
multi-shot sampling at :118,156,183).  ``Batch`` plans every instance's gate
list into LDS tile passes, hands them to the engine's batch object
(qvmhip_batch_*), which schedules them level by level -- one batched tile
kernel per level and tile width for all instances -- then computes each
instance's MaxCut cost <H_C> = sum_edges <Z_a Z_b> and draws its shots, all
replayed as one hipGraph.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _native as N
from . import circuits as CI
from . import gates as G
from . import planner as P


def tile_width(n: int, dtype: str, cap: int = 0) -> int:
    return min(n, cap or P.default_tile_m(dtype))


class Batch:
    def __init__(self, circuits: Sequence[Sequence], nqubits: Sequence[int], dtype: str = "f64",
                 edges: Sequence | None = None, shots: int = 0, uniforms=None, use_graph: bool = True,
                 fma_tiles: bool = False, tile_m: int = 0):
        """circuits[i]: list of (U, quil_qubits) on nqubits[i] qubits
        (every n >= 8); edges[i]: MaxCut edges for <H_C>; uniforms:
        count*shots draws in [0,1], instance-major."""
        self.count = len(circuits)
        self.nq = [int(x) for x in nqubits]
        self.dtype = dtype
        self.shots = int(shots)
        self._h = C.c_void_p()
        nqa = N.u32arr(self.nq)
        N.check(N.lib().qvmhip_batch_create(self.count, N.u32ptr(nqa), N.F64 if dtype == "f64" else N.F32,
                                            C.byref(self._h)))
        if fma_tiles:
            N.check(N.lib().qvmhip_batch_set_flags(self._h, N.FLAG_FMA_TILES))
        self.items = 0
        for i, gates in enumerate(circuits):
            n = self.nq[i]
            if n < 8:
                raise ValueError("batch instances need at least 8 qubits (tile passes)")
            m = tile_width(n, dtype, tile_m)
            plan = P.plan_tiled(list(gates), n, m, P.default_low(dtype) if m > 8 else 3)
            for it in plan:
                if it.kind == "tile":
                    tq = N.u32arr(it.qubits)
                    arr = (N.TileOp * len(it.ops))(*it.ops)
                    N.check(N.lib().qvmhip_batch_add_tile_pass(self._h, i, len(tq), N.u32ptr(tq), len(it.ops), arr))
                else:
                    U = N.cmat(it.U)
                    q = N.u32arr(it.qubits)
                    N.check(N.lib().qvmhip_batch_add_matrix(self._h, i, N.dptr(U.view(np.float64)), len(q),
                                                            N.u32ptr(q)))
                self.items += 1
        if edges is not None:
            for i, e in enumerate(edges):
                ea = N.u32arr(np.asarray(e, dtype=np.uint32).reshape(-1))
                N.check(N.lib().qvmhip_batch_set_zz(self._h, i, N.u32ptr(ea), len(ea) // 2))
        if self.shots:
            u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
            assert u.size == self.count * self.shots
            N.check(N.lib().qvmhip_batch_set_shots(self._h, self.shots, N.dptr(u)))
        N.check(N.lib().qvmhip_batch_finalize(self._h, 1 if use_graph else 0))

    def run(self) -> None:
        N.check(N.lib().qvmhip_batch_run(self._h))

    def sync(self) -> None:
        N.check(N.lib().qvmhip_batch_sync(self._h))

    def results(self):
        e = np.zeros(self.count, dtype=np.float64)
        s = np.zeros(max(1, self.count * self.shots), dtype=np.uint64)
        N.check(N.lib().qvmhip_batch_results(self._h, N.dptr(e), N.u64ptr(s)))
        return e, s[: self.count * self.shots].reshape(self.count, self.shots)

    def amplitudes(self, inst: int) -> np.ndarray:
        n = self.nq[inst]
        out = np.empty(1 << n, dtype=np.complex128)
        N.check(N.lib().qvmhip_batch_read_amplitudes(self._h, inst, 0, 1 << n, N.dptr(out.view(np.float64))))
        return out

    def close(self) -> None:
        if self._h and self._h.value:
            N.lib().qvmhip_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def qaoa_batch(instances, dtype: str = "f64", shots: int = 1024, seed: int = 9, use_graph: bool = True,
               fma_tiles: bool = False, tile_m: int = 0):
    """Batch for (n, edges, gammas, betas) instances (circuits.qaoa_instances)."""
    circuits = [CI.qaoa_maxcut(n, e, g, b) for n, e, g, b in instances]
    u = CI.uniforms(seed, len(instances) * shots) if shots else None
    return Batch(circuits, [n for n, _, _, _ in instances], dtype, edges=[e for _, e, _, _ in instances],
                 shots=shots, uniforms=u, use_graph=use_graph, fma_tiles=fma_tiles, tile_m=tile_m), circuits, u


def expected_cut(nedges: int, hc: float) -> float:
    """(|E| - <H_C>) / 2 (examples/qaoa.lisp:68-75)."""
    return 0.5 * (nedges - hc)
