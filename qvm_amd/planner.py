"""Execution planners: turn a gate list into engine launches.

Three plans, all consumed by ``compile_gates``:

* ``plain``  -- one engine launch per gate (apply-matrix-operator per gate,
  src/wavefunction.lisp:308-330).  The reference the other two are checked
  against.
* ``tiled``  -- gates grouped into tile passes: a pass names up to m tile
  qubits and applies every gate that fits, in dependency order, to each
  2^m-amplitude block while it sits in registers.  Per-gate arithmetic is the
  plain kernels', so FP64 results are bit-identical to ``plain`` (unless
  ``merge_1q`` multiplies runs of 1-qubit gates together first).
* ``fused``  -- gates multiplied into dense 2^k x 2^k blocks (k <= kmax) that
  the MFMA kernel applies in one sweep each; the fusion of
  compile-loaded-program with *fuse-gates-during-compilation*
  (src/qvm.lisp:166-175).  Rounding differs from plain at the 1e-16 level.

Both grouping planners use the same dependency rule: a gate may join the
current group once every earlier gate that shares a qubit with it has been
scheduled, so gates on disjoint qubits are reordered freely and the result is
the same unitary.

Gates are (U, quil_qubits) pairs: U in the Quil convention (first argument is
the most significant matrix bit).
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from . import gates as G

Gate = Tuple[np.ndarray, Sequence[int]]


# ---------------------------------------------------------------------------
# Dependency-ordered grouping


class _Ready:
    """Tracks which gates are ready: all earlier gates on their qubits done."""

    def __init__(self, gates_qubits: List[Tuple[int, ...]], nq: int):
        self.qs = gates_qubits
        self.queues = [deque() for _ in range(nq)]
        for i, qs in enumerate(gates_qubits):
            for q in qs:
                self.queues[q].append(i)
        self.done = [False] * len(gates_qubits)
        self.remaining = len(gates_qubits)

    def is_ready(self, i: int) -> bool:
        return all(self.queues[q] and self.queues[q][0] == i for q in self.qs[i])

    def ready(self) -> List[int]:
        heads = set()
        for dq in self.queues:
            if dq:
                heads.add(dq[0])
        return sorted(i for i in heads if self.is_ready(i))

    def take(self, i: int) -> None:
        for q in self.qs[i]:
            assert self.queues[q][0] == i
            self.queues[q].popleft()
        self.done[i] = True
        self.remaining -= 1


def group_in_order(gates_qubits: List[Tuple[int, ...]], width: int, seed_qubits=(),
                   standalone=lambda i: False, anywhere=lambda i: False) -> List[Tuple[str, List[int], List[int]]]:
    """Groups of consecutive gates (original order kept, so the grouped
    execution performs exactly the plain path's floating-point operations).
    ``anywhere(i)``: gate i joins the current group whatever its qubits (a
    diagonal the tile kernel applies from the block's base index)."""
    out = []
    T: set = set(seed_qubits)
    members: List[int] = []
    for i, qs in enumerate(gates_qubits):
        if standalone(i):
            if members:
                out.append(("group", sorted(T), members))
            out.append(("single", list(qs), [i]))
            T, members = set(seed_qubits), []
            continue
        if anywhere(i):
            members.append(i)
            continue
        if len(T | set(qs)) > width:
            out.append(("group", sorted(T), members))
            T, members = set(seed_qubits), []
        T |= set(qs)
        members.append(i)
    if members:
        out.append(("group", sorted(T), members))
    return out


def group_gates(gates_qubits: List[Tuple[int, ...]], nq: int, width: int, seed_qubits=(),
                standalone=lambda i: False, anywhere=lambda i: False) -> List[Tuple[str, List[int], List[int]]]:
    """Greedy dependency-aware grouping.

    Returns a list of (kind, qubit set, gate indices in application order);
    kind is "group" or "single" (a gate that ``standalone`` excludes from
    groups, applied on its own).  Gates on disjoint qubits may move past each
    other: the unitary is unchanged, the rounding of the result is not
    (floating-point gate applications on disjoint qubits do not commute
    bit-for-bit).  A gate with ``anywhere(i)`` joins a group as soon as it is
    ready, without adding its qubits to the group's set.
    """
    R = _Ready(gates_qubits, nq)
    out = []
    while R.remaining:
        ready = R.ready()
        single = [i for i in ready if standalone(i)]
        if single and all(standalone(i) for i in ready):
            i = single[0]
            R.take(i)
            out.append(("single", list(gates_qubits[i]), [i]))
            continue
        T = set(q for q in seed_qubits if q < nq)
        members: List[int] = []
        while True:
            progressed = False
            for i in R.ready():
                if standalone(i):
                    continue
                if anywhere(i) or set(gates_qubits[i]) <= T:
                    R.take(i)
                    members.append(i)
                    progressed = True
            if progressed:
                continue
            # extend the tile with the qubits of the earliest ready gate that fits
            best = None
            for i in R.ready():
                if standalone(i):
                    continue
                new = set(gates_qubits[i]) - T
                if len(T) + len(new) <= width:
                    if best is None or len(new) < len(set(gates_qubits[best]) - T):
                        best = i
            if best is None:
                break
            T |= set(gates_qubits[best])
        if not members:
            # nothing fits an empty group: take the earliest ready gate alone
            i = R.ready()[0]
            R.take(i)
            out.append(("single", list(gates_qubits[i]), [i]))
            continue
        out.append(("group", sorted(T), members))
    return out


# ---------------------------------------------------------------------------
# Tile ops


def _is_diag(U):
    return np.count_nonzero(U - np.diag(np.diag(U))) == 0


def _gdiag_op(U: np.ndarray, nat: Sequence[int], local) -> N.TileOp:
    """A diagonal gate with qubits outside the tile as a GDIAG op: local
    qubit q0 (any local position when the gate has none inside) and the
    outside qubits g0, g1; entry index b(q0) + 2 b(g0) + 4 b(g1)."""
    d = np.diag(U)
    inside = [q for q in nat if q in local]
    outside = [q for q in nat if q not in local]
    if len(inside) > 1 or not 1 <= len(outside) <= 2:
        raise ValueError("GDIAG covers one tile qubit and one or two outside qubits")
    q0 = local[inside[0]] if inside else 0
    ng = len(outside)
    vals = np.zeros(1 << (1 + ng), dtype=np.complex128)
    for e in range(vals.size):
        j = 0  # the gate's diagonal index: bit b <-> nat[b]
        for b, q in enumerate(nat):
            bit = (e & 1) if q not in outside else (e >> (1 + outside.index(q))) & 1
            j |= bit << b
        vals[e] = d[j]
    op = N.TileOp()
    op.kind = N.TILE_GDIAG
    op.q0 = q0
    op.q1 = outside[0] | ((outside[1] << 8) if ng == 2 else 0) | (ng << 16)
    for i, x in enumerate(vals.view(np.float64)):
        op.m[i] = x
    return op


def tile_op(U: np.ndarray, quil_qubits: Sequence[int], local) -> N.TileOp:
    """Translate one 1- or 2-qubit gate into a tile op on local positions
    (a diagonal gate may also name qubits outside the tile: GDIAG)."""
    op = N.TileOp()
    m = np.zeros(32)
    nat = G.nat_tuple(*quil_qubits)
    k = len(nat)
    if any(q not in local for q in nat):
        if not _is_diag(U):
            raise ValueError("only diagonal gates may name qubits outside the tile")
        return _gdiag_op(U, nat, local)
    if k == 1:
        op.q0 = local[nat[0]]
        if _is_diag(U):
            op.kind = N.TILE_DIAG1
            d = np.diag(U)
            m[0:4] = [d[0].real, d[0].imag, d[1].real, d[1].imag]
        else:
            op.kind = N.TILE_DENSE1
            m[0:8] = np.ascontiguousarray(U).reshape(-1).view(np.float64)
    elif k == 2:
        op.q0, op.q1 = local[nat[0]], local[nat[1]]
        I4 = np.eye(4)
        d = np.diag(U)
        if _is_diag(U) and d[0] == 1 and d[1] == 1 and d[2] == 1:
            op.kind = N.TILE_CPHASE
            m[0:2] = [d[3].real, d[3].imag]
        elif _is_diag(U):
            op.kind = N.TILE_DIAG2
            m[0:8] = np.ascontiguousarray(d).view(np.float64)
        elif np.array_equal(U, G.gate_matrix("CNOT")):
            op.kind = N.TILE_CPERM  # q0 = target (nat[0]), q1 = control (nat[1])
        elif np.array_equal(U, G.gate_matrix("SWAP")):
            op.kind = N.TILE_SWAP
        elif (np.array_equal(U[:2, :2], I4[:2, :2]) and not U[:2, 2:].any() and not U[2:, :2].any()):
            # CONTROLLED 1q: control = first Quil argument = nat[1]
            op.kind = N.TILE_DENSE1
            op.q0 = local[nat[0]]
            op.q1 = 0
            op.ctrl_mask = 1 << local[nat[1]]
            m[0:8] = np.ascontiguousarray(U[2:, 2:]).reshape(-1).view(np.float64)
        else:
            op.kind = N.TILE_DENSE2
            m[0:32] = np.ascontiguousarray(U).reshape(-1).view(np.float64)
    else:
        raise ValueError("tile ops are 1- or 2-qubit gates")
    for i in range(32):
        op.m[i] = m[i]
    return op


# ---------------------------------------------------------------------------
# Dense block fusion


def embed(U: np.ndarray, quil_qubits: Sequence[int], block: Sequence[int]) -> np.ndarray:
    """Matrix of U on ``block`` (index bit j <-> block[j])."""
    nat = G.nat_tuple(*quil_qubits)
    kb = len(block)
    pos = [block.index(q) for q in nat]
    D = 1 << kb
    idx = np.arange(D)
    gin = np.zeros(D, dtype=np.int64)
    for j, p in enumerate(pos):
        gin |= ((idx >> p) & 1) << j
    rest = idx.copy()
    for p in pos:
        rest &= ~(1 << p)
    E = np.zeros((D, D), dtype=np.complex128)
    kg = len(nat)
    for r in range(1 << kg):
        dep = 0
        for j, p in enumerate(pos):
            if (r >> j) & 1:
                dep |= 1 << p
        E[rest | dep, idx] = U[r, gin]
    return E


def fuse_groups(gates: List[Gate], nq: int, kmax: int):
    """Dense blocks: list of (U_block, block_qubits) in engine order
    (block_qubits[j] = index bit j)."""
    qs = [tuple(q) for _, q in gates]
    groups = group_gates(qs, nq, kmax, standalone=lambda i: len(qs[i]) > kmax)
    blocks = []
    for kind, T, members in groups:
        block = list(T)
        M = np.eye(1 << len(block), dtype=np.complex128)
        for i in members:
            U, q = gates[i]
            M = embed(U, q, block) @ M
        blocks.append((M, block, members))
    return blocks


# ---------------------------------------------------------------------------
# Plans


@dataclass
class PlanItem:
    kind: str                 # "gate" | "tile" | "fused"
    qubits: List[int]         # engine order (gate/fused) or tile qubits
    U: Optional[np.ndarray] = None
    ops: list = field(default_factory=list)
    gates: List[int] = field(default_factory=list)


def plan_plain(gates: List[Gate], nq: int) -> List[PlanItem]:
    return [PlanItem("gate", list(G.nat_tuple(*q)), U=U, gates=[i]) for i, (U, q) in enumerate(gates)]


def _kind_1q(U: np.ndarray) -> str:
    """How the tile kernel runs a 2x2 gate: "diag" (a diagonal op, free on a
    register bit in FMA passes), "reduced" (real entries, or a real diagonal
    with an imaginary anti-diagonal: half the multiplies) or "general"."""
    if U[0, 1] == 0 and U[1, 0] == 0:
        return "diag"
    if not np.any(U.imag):
        return "reduced"
    if not np.any(U[[0, 1], [0, 1]].imag) and not np.any(U[[0, 1], [1, 0]].real):
        return "reduced"
    return "general"


# Relative cost of one tile op per kind in FMA passes, from per-op timings at
# n = 30 (FP64: general U3 0.27-0.29 ms, RX 0.12-0.14, RZ 0 on a register bit
# and 0.19 on a lane bit).
_COST_1Q = {"diag": 0.8, "reduced": 1.0, "general": 2.2}


def _merge_runs_1q(gates: List[Gate]) -> List[Gate]:
    """Multiply consecutive 1-qubit gates on the same qubit (in dependency
    order) into one matrix when that is not dearer for the tile kernel than
    keeping them apart: RX then RZ stay two ops (a reduced op and a
    diagonal, which the FMA scheduler folds into the register diagonals)
    instead of one general op, while runs of general gates become one.
    Measured at n = 30 (FMA passes, ms): layered-40 FP64 759.7 merged / 728.2
    kept apart, FP32 442.8 / 398.2; random 60 % U3 / 40 % CNOT-CZ FP64 230.0 /
    286.5, FP32 121.8 / 159.6."""
    # pending[q] = index in `merged` of the open 1-qubit run on q; a gate on
    # q with more qubits closes the run.
    merged: List[Gate] = []
    pending = {}
    for U, q in gates:
        q = list(q)
        if len(q) == 1:
            j = pending.get(q[0])
            if j is not None:
                V = U @ merged[j][0]
                if _COST_1Q[_kind_1q(V)] <= _COST_1Q[_kind_1q(merged[j][0])] + _COST_1Q[_kind_1q(U)]:
                    merged[j] = (V, merged[j][1])
                    continue
            merged.append((U, q))
            pending[q[0]] = len(merged) - 1
        else:
            for x in q:
                pending.pop(x, None)
            merged.append((U, q))
    return merged


def plan_tiled(gates: List[Gate], nq: int, m: int, low: int = 4, merge_1q: bool = False,
               reorder: bool = True, wide: str = "gate") -> List[PlanItem]:
    """Register-tile passes of up to m qubits that always include qubits
    0..low-1 (low sets how many low-order qubits every block keeps as one
    contiguous run; the other tile qubits are chosen per pass).

    reorder=False keeps the original gate order, so the FP64 result is
    bit-identical to the plain path; reorder=True lets commuting gates move
    into earlier passes (far fewer passes, results equal to rounding).
    Gates on more than 2 qubits run outside the passes: as plain gates
    (wide="gate") or, for 3..6 qubits, as dense blocks on the matrix cores
    (wide="fused")."""
    if merge_1q:
        gates = _merge_runs_1q(gates)
    if nq < 8:
        return plan_plain(gates, nq)  # tile passes need at least 8 qubits
    m = min(m, nq)
    low = min(low, m)
    qs = [tuple(q) for _, q in gates]
    # 1- and 2-qubit diagonals run in any pass (GDIAG reads the qubits
    # outside the tile from the block's base index)
    diag = [len(q) <= 2 and _is_diag(U) for U, q in gates]
    if reorder:
        groups = group_gates(qs, nq, m, seed_qubits=range(low), standalone=lambda i: len(qs[i]) > 2,
                             anywhere=lambda i: diag[i])
    else:
        groups = group_in_order(qs, m, seed_qubits=range(low), standalone=lambda i: len(qs[i]) > 2,
                                anywhere=lambda i: diag[i])
    plan = []
    for kind, T, members in groups:
        if kind == "single":
            i = members[0]
            U, q = gates[i]
            if wide == "fused" and 3 <= len(q) <= 6:
                # the fused kernel's convention: block[j] = matrix index bit j
                plan.append(PlanItem("fused", list(G.nat_tuple(*q)), U=U, gates=[i]))
            else:
                plan.append(PlanItem("gate", list(G.nat_tuple(*q)), U=U, gates=[i]))
            continue
        # pad the tile to m qubits with the lowest unused qubits (contiguity)
        T = set(T)
        q = 0
        while len(T) < m:
            if q not in T:
                T.add(q)
            q += 1
        tq = sorted(T)
        local = {x: j for j, x in enumerate(tq)}
        ops = [tile_op(gates[i][0], gates[i][1], local) for i in members]
        plan.append(PlanItem("tile", tq, ops=ops, gates=members))
    return plan


def plan_fused(gates: List[Gate], nq: int, kmax: int) -> List[PlanItem]:
    plan = []
    for M, block, members in fuse_groups(gates, nq, kmax):
        if len(block) >= 3:
            plan.append(PlanItem("fused", block, U=M, gates=members))
        else:
            # engine order: block[j] = index bit j
            plan.append(PlanItem("gate", block, U=M, gates=members))
    return plan


def execute_plan(state, plan: List[PlanItem]) -> None:
    """Run a plan eagerly (one C-ABI call per item)."""
    for it in plan:
        if it.kind == "gate":
            state.apply_matrix(it.U, it.qubits)
        elif it.kind == "fused":
            state.apply_fused(it.U, it.qubits)
        else:
            state.apply_tile_pass(it.qubits, it.ops)


def build_program(state, plan: List[PlanItem], use_graph: bool = True):
    from .state import Program
    p = Program(state)
    for it in plan:
        if it.kind == "gate":
            p.add_matrix(it.U, it.qubits)
        elif it.kind == "fused":
            p.add_fused(it.U, it.qubits)
        else:
            p.add_tile_pass(it.qubits, it.ops)
    p.finalize(use_graph=use_graph)
    return p


def default_tile_m(dtype: str) -> int:
    return 11 if dtype == "f64" else 12


def default_low(dtype: str) -> int:
    """Low-order qubits every pass keeps as one contiguous run: one 128 B line
    (8 FP64 / 16 FP32 amplitudes).  Measured at n = 30 (layered-40, QFT-30,
    Haar-U3 + CZ depth 20, reordered FMA passes): FP64 runs of 8 took 750.7 /
    161.5 / 391.6 ms against 778.9 / 166.0 / 403.3 for 16 (and in-order exact
    layered passes 1942 against 2145 ms); FP32 runs of 16 took 442.7 / 89.5 /
    236.1 against 448.3 / 94.4 / 239.2 for 32.  A longer run leaves fewer tile
    qubits free for the other targets, so a circuit needs more passes."""
    return 3 if dtype == "f64" else 4


def max_tile_m(dtype: str) -> int:
    """Widest tile the engine runs (FP64 12, FP32 13 qubits)."""
    return 12 if dtype == "f64" else 13


def compile_gates(state, gates: List[Gate], mode: str = "tiled", use_graph: bool = True, m: int = 0,
                  kmax: int = 4, merge_1q: bool = False, reorder: bool = True):
    from .qvm import CompiledProgram
    nq = state.num_qubits
    if mode == "plain":
        plan = plan_plain(gates, nq)
    elif mode == "tiled":
        plan = plan_tiled(gates, nq, m or default_tile_m(state.dtype), default_low(state.dtype), merge_1q,
                          reorder)
    elif mode == "fused":
        plan = plan_fused(gates, nq, kmax)
    elif mode == "auto":
        # the candidate plan the trained launch-cost model prices lowest
        from . import costmodel as CM
        cm = CM.CostModel.shared()
        _, plan, _ = cm.choose(gates, nq, state.dtype)
    else:
        raise ValueError(f"unknown plan mode {mode!r}")
    return CompiledProgram(build_program(state, plan, use_graph), plan, mode)


def sweeps(plan: List[PlanItem]) -> int:
    return len(plan)
