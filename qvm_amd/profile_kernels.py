"""Short kernel workloads for hardware-counter profiling (rocprofv3 --pmc).

Runs `reps` launches of one kernel family on an n-qubit state, then one
device-to-device copy of the whole state (a transfer of known size:
2^n * B bytes read and written) so counter readings can be scaled against
it.  Usage:

    python -m qvm_amd.profile_kernels --n 30 --kernel dense1 --reps 8
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
                    choices=["dense1", "diag_full", "perm", "diag_sparse", "fused3", "fused4", "fused5", "fused6",
                             "tile"])
    ap.add_argument("--reps", type=int, default=8)
    ap.add_argument("--fma", action="store_true", help="tile passes with FMA-contracted products")
    ap.add_argument("--tile-m", type=int, default=0, help="tile width (default: the planner's)")
    ap.add_argument("--qubits", default="",
                    help="comma-separated target qubits, cycled (fused blocks: the first k; default: random)")
    a = ap.parse_args(argv)
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


if __name__ == "__main__":
    sys.exit(main())
