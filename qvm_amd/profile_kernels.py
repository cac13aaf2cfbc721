"""Short kernel workloads for hardware-counter profiling (rocprofv3 --pmc).

Runs `reps` launches of one kernel family on an n-qubit state, then one
device-to-device copy of the whole state (a transfer of known size:
2^n * B bytes read and written) so counter readings can be scaled against
it.  Usage:

    python -m qvm_amd.profile_kernels --n 30 --kernel dense1 --reps 8

`--kernel tile_ops` times one tile pass of `--count` identical `--op` gates
on tile-local qubit `--local` (by default the top tile qubit, a register bit
after the scheduler's remap; low tile qubits are lane bits), `--reps` times,
and prints the mean time per pass: the difference between two counts is the
cost of one op beyond the pass's load and store.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import gates as G
from .state import HipPureState


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--dtype", default="f64")
    ap.add_argument("--kernel", default="dense1",
                    choices=["dense1", "diag_full", "perm", "perm_adj", "swap", "diag_sparse", "fused3", "fused4",
                             "fused5", "fused6", "tile", "tile_ops"])
    ap.add_argument("--reps", type=int, default=8)
    ap.add_argument("--fma", action="store_true", help="tile passes with FMA-contracted products")
    ap.add_argument("--tile-m", type=int, default=0, help="tile width (default: the planner's)")
    ap.add_argument("--qubits", default="",
                    help="comma-separated target qubits, cycled (fused blocks: the first k; default: random)")
    ap.add_argument("--op", default="RX", choices=["RX", "U3", "RZ"], help="tile_ops: the gate")
    ap.add_argument("--count", type=int, default=8, help="tile_ops: gates in the pass")
    ap.add_argument("--local", type=int, default=-1, help="tile_ops: tile-local target (default: m - 1)")
    a = ap.parse_args(argv)
    if a.kernel == "tile_ops":
        return tile_ops(a)
    fixed = [int(x) for x in a.qubits.split(",") if x.strip()]
    st = HipPureState(a.n, a.dtype, fma_tiles=a.fma)
    rng = np.random.default_rng(0)
    for r in range(a.reps):
        q = int(rng.integers(0, a.n))
        if fixed:
            q = fixed[r % len(fixed)]
        if a.kernel == "dense1":
            st.apply_matrix(G.gate_matrix("RX", [0.1 + r]), [q])
        elif a.kernel == "diag_full":
            st.apply_matrix(G.gate_matrix("RZ", [0.1 + r]), [q])
        elif a.kernel == "perm":
            st.apply_matrix(G.gate_matrix("CNOT"), [q, (q + 7) % a.n])
        elif a.kernel == "perm_adj":
            st.apply_matrix(G.gate_matrix("CNOT"), [q, (q + 1) % a.n])
        elif a.kernel == "swap":
            st.apply_matrix(G.gate_matrix("SWAP"), [q, a.n - 1 - q])
        elif a.kernel == "diag_sparse":
            st.apply_matrix(G.gate_matrix("CPHASE", [0.1 + r]), [q, (q + 7) % a.n])
        elif a.kernel.startswith("fused"):
            k = int(a.kernel[-1])
            Z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
            qs = fixed[:k] if len(fixed) >= k else [int(x) for x in rng.choice(a.n, k, replace=False)]
            st.apply_fused(np.linalg.qr(Z)[0], qs)
        elif a.kernel == "tile":
            from . import circuits as C
            from . import planner as P
            plan = P.plan_tiled(C.layered(a.n, 1, seed=r), a.n, a.tile_m or P.default_tile_m(a.dtype),
                                P.default_low(a.dtype), merge_1q=a.fma)
            P.execute_plan(st, plan)
    st.sync()
    other = HipPureState(a.n, a.dtype)
    other.copy_from(st)  # known-size transfer: 2^n * B read + 2^n * B written
    other.sync()
    print(f"profiled {a.reps} x {a.kernel} at n={a.n} {a.dtype}; copy of {(1 << a.n) * (16 if a.dtype == 'f64' else 8)} B",
          flush=True)
    other.close()
    st.close()
    return 0


def tile_ops(a) -> int:
    from . import planner as P
    st = HipPureState(a.n, a.dtype, fma_tiles=a.fma)
    m = a.tile_m or P.default_tile_m(a.dtype)
    q = a.local if a.local >= 0 else m - 1
    U = {"RX": G.gate_matrix("RX", [0.3]), "U3": G.gate_matrix("U3", [0.3, 0.7, 1.1]),
         "RZ": G.gate_matrix("RZ", [0.3])}[a.op]
    tq = list(range(m))
    local = {t: t for t in tq}
    ops = [P.tile_op(U, [q], local) for _ in range(a.count)]
    prog = P.build_program(st, [P.PlanItem("tile", tq, ops=ops)], use_graph=False)
    prog.run()
    st.sync()
    ms = prog.time(a.reps)
    print(f"tile pass of {a.count} x {a.op} on tile qubit {q}: m={m} n={a.n} {a.dtype} "
          f"{'fma' if a.fma else 'exact'}: {ms:.4f} ms per pass", flush=True)
    prog.close()
    st.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
