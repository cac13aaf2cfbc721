"""Many small circuits in one captured launch sequence (QAOA parameter sweeps).

The reference runs a QAOA sweep one QVM at a time (examples/qaoa.lisp:7-75,
multi-shot sampling at :118,156,183).  ``Batch`` plans every instance's gate
list into LDS tile passes, hands them to the engine's batch object
(qvmhip_batch_*), which schedules them level by level -- one batched tile
kernel per level and tile width for all instances -- then computes each
instance's MaxCut cost <H_C> = sum_edges <Z_a Z_b> and draws its shots, all
replayed as one hipGraph.  ``autotune_tile_m`` picks the tile width by timing
candidate widths on a sample of the instances.
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


def tile_candidates(dtype: str) -> List[int]:
    """Tile widths the autotuner tries: the planner's default and one either side
    (FP64 10..12, FP32 11..13; 13 is FP32-only)."""
    m = P.default_tile_m(dtype)
    return [x for x in (m - 1, m, m + 1) if 8 <= x <= (12 if dtype == "f64" else 13)]


def autotune_tile_m(circuits: Sequence[Sequence], nqubits: Sequence[int], dtype: str = "f64",
                    fma_tiles: bool = True, candidates: Sequence[int] | None = None, sample_every: int = 4,
                    reps: int = 2):
    """Pick the batch tile width by measurement: build the batch of every
    ``sample_every``-th instance once per candidate width, time ``reps``
    graph launches of each (after one warm-up) with the engine's own kernels,
    and return (best width, {width: seconds per run}).  The sample keeps the
    mix of qubit counts; a width above an instance's qubit count tiles that
    instance whole, as the full batch will."""
    import time
    cands = list(candidates) if candidates else tile_candidates(dtype)
    idx = list(range(0, len(circuits), max(1, int(sample_every))))
    sc = [circuits[i] for i in idx]
    sn = [nqubits[i] for i in idx]
    timings = {}
    for m in cands:
        b = Batch(sc, sn, dtype, use_graph=True, fma_tiles=fma_tiles, tile_m=m)
        try:
            b.run()
            b.sync()
            t0 = time.perf_counter()
            for _ in range(max(1, reps)):
                b.run()
            b.sync()
            timings[m] = (time.perf_counter() - t0) / max(1, reps)
        finally:
            b.close()
    best = min(timings, key=timings.get)
    return best, timings


def qaoa_batch(instances, dtype: str = "f64", shots: int = 1024, seed: int = 9, use_graph: bool = True,
               fma_tiles: bool = False, tile_m=0):
    """Batch for (n, edges, gammas, betas) instances (circuits.qaoa_instances).
    ``tile_m``: 0 for the planner's default width, a width, or "auto" for
    the width ``autotune_tile_m`` measures fastest on a sample of the
    instances (the Batch then carries ``tile_m`` and ``tile_timings``)."""
    circuits = [CI.qaoa_maxcut(n, e, g, b) for n, e, g, b in instances]
    nq = [n for n, _, _, _ in instances]
    timings = None
    if tile_m == "auto":
        tile_m, timings = autotune_tile_m(circuits, nq, dtype, fma_tiles=fma_tiles)
    u = CI.uniforms(seed, len(instances) * shots) if shots else None
    b = Batch(circuits, nq, dtype, edges=[e for _, e, _, _ in instances],
              shots=shots, uniforms=u, use_graph=use_graph, fma_tiles=fma_tiles, tile_m=int(tile_m))
    b.tile_m = int(tile_m) or P.default_tile_m(dtype)
    b.tile_timings = timings
    return b, circuits, u


def expected_cut(nedges: int, hc: float) -> float:
    """(|E| - <H_C>) / 2 (examples/qaoa.lisp:68-75)."""
    return 0.5 * (nedges - hc)
