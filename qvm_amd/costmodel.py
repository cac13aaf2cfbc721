"""Learned plan cost model.

A small MLP predicts the duration of one planned launch (a plain gate, a
register-tile pass or a fused MFMA block) from features of the launch.  The
planner prices every candidate plan for a circuit by summing the
predictions, and runs the cheapest.  The candidates span the fusion width
(plain; dense blocks of k = 2..6 qubits on the matrix cores), the tile width
m and the qubit permutation into the on-chip block (the number of low-order
qubits every tile keeps contiguous; the other tile qubits follow the
gates), and a hybrid (tile passes for 1- and 2-qubit gates, MFMA blocks for
the wider ones).  The network is trained and evaluated by the engine's own
HIP kernels (qvm_amd/csrc/mlp.hip) on timings the engine records with HIP
events (``collect``) on states whose tile passes use the FMA form (the fast
mode the planner chooses for).

Two planners without the network are kept for the ablation: the fixed rule
(tile passes at the default width and low run, the planner's "tiled" mode)
and an analytic model (launch overhead + HBM bytes at a streaming rate +
instruction issue at the chip's vector rate, or MFMA flops for fused
blocks).

The target is log(ms per GB of state): launches of any n share one scale, and
the model's prediction for n is exp(y) * state_GB.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from typing import Dict, List, Sequence

import numpy as np

from . import _native as N
from . import circuits as CI
from . import gates as G
from . import planner as P

DIN = 18
HIDDEN = 32

# Scheduled tile-op kinds (the engine's slot-space program, tile.hip).
_DENSE_KINDS = {N.TILE_DENSE1, N.TILE_DENSE2, N.TILE_XREG, N.TILE_D1REAL, N.TILE_D1RX}
_DIAG_KINDS = {N.TILE_DIAG1, N.TILE_DIAG2, N.TILE_CPHASE, N.TILE_GDIAG, N.TILE_RDIAG, N.TILE_CUT_PHASE}
_REMAP = N.TILE_REMAP


def tile_profile(item: "P.PlanItem", dtype: str):
    """(scheduled ops, remaps, dense ops, diagonal ops) of a tile pass as the
    engine's scheduler lays it out for an FMA state (cached on the item)."""
    prof = getattr(item, "_prof", None)
    if prof is None:
        sched = N.tile_schedule(dtype, len(item.qubits), item.ops, fma=True) if item.ops else []
        kinds = [op.kind for op in sched]
        prof = (len(kinds), sum(k == _REMAP for k in kinds), sum(k in _DENSE_KINDS for k in kinds),
                sum(k in _DIAG_KINDS for k in kinds))
        item._prof = prof
    return prof


def low_run(qs) -> int:
    """Length of the run of qubits 0, 1, 2, ... in qs (contiguous amplitudes)."""
    s, L = set(qs), 0
    while L in s:
        L += 1
    return L


def launch_features(item: "P.PlanItem", n: int, dtype: str) -> np.ndarray:
    """Features of one launch (all in [0, ~1])."""
    f = np.zeros(DIN, dtype=np.float32)
    qs = list(item.qubits)
    kind = item.kind
    f[0] = 1.0
    f[1] = kind == "gate"
    f[2] = kind == "tile"
    f[3] = kind == "fused"
    f[4] = len(qs) / 13.0
    if kind == "tile":
        ns, nr, nd, ng = tile_profile(item, dtype)
        f[5] = min(ns, 256) / 128.0
        f[6] = min(nr, 64) / 32.0
        f[7] = min(nd, 128) / 64.0
        f[8] = min(ng, 128) / 64.0
    f[9] = min(qs) / 34.0
    f[10] = max(qs) / 34.0
    f[11] = low_run(qs) / 8.0
    if kind == "gate":
        U = item.U
        k = G.num_qubits_of(U)
        f[12] = G.changed_amplitudes(U, k) / float(1 << k)
        f[13] = 1.0 if np.count_nonzero(U - np.diag(np.diag(U))) == 0 else 0.0
    else:
        f[12] = 1.0
    if kind == "fused":
        f[14] = (1 << len(qs)) / 64.0
    f[15] = 1.0 if dtype == "f32" else 0.0
    f[16] = n / 34.0
    f[17] = max(0.0, (24 - n) / 8.0)
    return f


def state_gb(n: int, dtype: str) -> float:
    return (1 << n) * (16 if dtype == "f64" else 8) / 1e9


class AnalyticModel:
    """Launch time from first principles, no training: a fixed launch
    overhead, plus HBM bytes (read and write of every amplitude the launch
    touches) at a streaming rate, plus -- for tile passes -- vector
    instruction issue (lane-ops per amplitude per scheduled op at the chip's
    FP64 / FP32 vector rate, scaled by the share of the chip the pass's
    wavefronts can fill) overlapped with the transfer, or -- for fused blocks
    -- 8 * 2^k flops per amplitude at 3/4 of the matrix-core peak."""

    OVERHEAD_S = 5e-6
    STREAM_BPS = 6.0e12                      # measured in-place sweep rate, MI355X
    LANE_OPS_PER_S = {"f64": 3.93e13, "f32": 7.86e13}
    MFMA_FLOPS = {"f64": 78.6e12, "f32": 157.3e12}
    OP_COST = {"remap": 16.0, "dense": 12.0, "diag": 10.0, "other": 4.0}  # lane-ops per amplitude
    WAVE_SLOTS = 256 * 4 * 4                 # CUs x SIMDs x resident waves per SIMD

    def seconds(self, kind: str, n: int, dtype: str, width: int, changed: int = 0, prof=(0, 0, 0, 0)) -> float:
        """kind "gate" (changed = amplitudes it can change), "fused" (width
        = k) or "tile" (prof = tile_profile of the pass)."""
        amp = 16 if dtype == "f64" else 8
        if kind == "gate":
            return self.OVERHEAD_S + 2 * changed * amp / self.STREAM_BPS
        mem = 2 * (1 << n) * amp / self.STREAM_BPS
        if kind == "fused":
            flops = 8.0 * (1 << width) * (1 << n)
            return self.OVERHEAD_S + max(mem, flops / (0.75 * self.MFMA_FLOPS[dtype]))
        ns, nr, nd, ng = prof
        lane_ops = (nr * self.OP_COST["remap"] + nd * self.OP_COST["dense"] + ng * self.OP_COST["diag"]
                    + (ns - nr - nd - ng) * self.OP_COST["other"]) * (1 << n)
        regs = 8 if dtype == "f64" else 16
        waves = max(1.0, (1 << n) / (64.0 * regs))
        share = min(1.0, waves / self.WAVE_SLOTS)
        issue = lane_ops / (self.LANE_OPS_PER_S[dtype] * share)
        return self.OVERHEAD_S + max(mem, issue) + 0.25 * min(mem, issue)

    def item_seconds(self, item: "P.PlanItem", n: int, dtype: str) -> float:
        if item.kind == "gate":
            k = G.num_qubits_of(item.U)
            return self.seconds("gate", n, dtype, k, changed=G.changed_amplitudes(item.U, k) << (n - k))
        if item.kind == "fused":
            return self.seconds("fused", n, dtype, len(item.qubits))
        return self.seconds("tile", n, dtype, len(item.qubits), prof=tile_profile(item, dtype))

    def record_seconds(self, r: Dict) -> float:
        """The model's time for a record of ``collect``."""
        return self.seconds(r["kind"], r["n"], r["dtype"], r["width"], r.get("changed", 0),
                            tuple(r.get("prof", (0, 0, 0, 0))))

    def predict_ms(self, items: List["P.PlanItem"], n: int, dtype: str) -> np.ndarray:
        return np.array([1e3 * self.item_seconds(it, n, dtype) for it in items])

    def plan_cost(self, plan: List["P.PlanItem"], n: int, dtype: str) -> float:
        return float(self.predict_ms(plan, n, dtype).sum())

    def choose(self, gates, n: int, dtype: str, candidates: Dict[str, List["P.PlanItem"]] | None = None):
        cands = candidates or candidate_plans(gates, n, dtype)
        costs = {name: self.plan_cost(plan, n, dtype) for name, plan in cands.items()}
        best = min(costs, key=costs.get)
        return best, cands[best], costs


class MLP:
    """Handle on the engine's MLP (weights live on the GPU)."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0):
        self._h = C.c_void_p()
        if params is None:
            params = init_params(seed)
        params = np.ascontiguousarray(params, dtype=np.float32)
        N.check(N.lib().qvmhip_mlp_create(DIN, HIDDEN, N.fptr(params), C.byref(self._h)))
        npar = C.c_int()
        N.check(N.lib().qvmhip_mlp_num_params(self._h, C.byref(npar)))
        self.num_params = npar.value

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, dtype=np.float32)
        N.check(N.lib().qvmhip_mlp_get_params(self._h, N.fptr(out)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.check(N.lib().qvmhip_mlp_set_params(self._h, N.fptr(p)))

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty(X.shape[0], dtype=np.float32)
        N.check(N.lib().qvmhip_mlp_predict(self._h, N.fptr(X), X.shape[0], N.fptr(out)))
        return out

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.num_params, dtype=np.float32)
        loss = C.c_float()
        N.check(N.lib().qvmhip_mlp_gradient(self._h, N.fptr(X), N.fptr(T), X.shape[0], N.fptr(g), C.byref(loss)))
        return loss.value, g

    def train(self, X, T, epochs: int = 2000, lr: float = 3e-3) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float()
        N.check(N.lib().qvmhip_mlp_train(self._h, N.fptr(X), N.fptr(T), X.shape[0], int(epochs), float(lr),
                                         C.byref(loss)))
        return loss.value

    def close(self) -> None:
        if self._h and self._h.value:
            N.lib().qvmhip_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation of the flat parameter vector."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, math.sqrt(2.0 / DIN), (HIDDEN, DIN))
    W2 = rng.normal(0, math.sqrt(2.0 / HIDDEN), (HIDDEN, HIDDEN))
    w3 = rng.normal(0, math.sqrt(1.0 / HIDDEN), HIDDEN)
    return np.concatenate([W1.ravel(), np.zeros(HIDDEN), W2.ravel(), np.zeros(HIDDEN), w3, [0.0]]).astype(np.float32)


# ---------------------------------------------------------------------------
# Timing collection: the engine records its own launch durations.


def _time_items(state, items: List["P.PlanItem"], reps: int = 3) -> float:
    prog = P.build_program(state, items, use_graph=False)
    prog.run()
    state.sync()
    ms = prog.time(reps)
    prog.close()
    return ms


def _family_circuit(n: int, rng: np.random.Generator) -> List:
    """A short random circuit on n qubits from one of the workload families
    (layered RX/RZ + CNOT, Haar 1q + CNOT/CZ, ZZ-phase heavy, dense 5- or
    6-qubit blocks, Haar U3 + CZ matchings)."""
    seed = int(rng.integers(1 << 30))
    fam = int(rng.integers(5))
    if fam == 0:
        return CI.layered(n, int(rng.integers(1, 4)), seed=seed)
    if fam == 1:
        return CI.random_mix(n, int(rng.integers(1, 4)), seed=seed)
    if fam == 2:
        return CI.diagonal_heavy(n, 1, seed=seed)
    if fam == 3:
        return CI.dense_blocks(n, int(rng.integers(1, 3)), int(rng.integers(5, 7)), seed=seed)
    return CI.haar_cz(n, int(rng.integers(1, 3)), seed=seed)


def random_launches(n: int, rng: np.random.Generator, count: int, dtype: str = "f64") -> List["P.PlanItem"]:
    """Launches drawn from the candidate plans themselves: a random family
    circuit is planned by a random candidate (plain, fused k = 2..6, tiled at
    one of the widths and low runs, hybrid) and a few of its launches are
    kept, so the training launches follow the distribution the planner
    prices."""
    items: List["P.PlanItem"] = []
    while len(items) < count:
        cands = candidate_plans(_family_circuit(n, rng), n, dtype)
        name = list(cands)[int(rng.integers(len(cands)))]
        plan = cands[name]
        take = min(len(plan), int(rng.integers(1, 5)), count - len(items))
        for j in rng.choice(len(plan), take, replace=False):
            items.append(plan[int(j)])
    return items


def collect(ns: Sequence[int], dtypes: Sequence[str], per_config: int, seed: int = 0) -> List[Dict]:
    """Time random launches at each (n, dtype); returns records with
    features and measured ms."""
    from .state import HipPureState
    rng = np.random.default_rng(seed)
    recs = []
    for n in ns:
        for dt in dtypes:
            st = HipPureState(n, dt, fma_tiles=True)
            for it in random_launches(n, rng, per_config, dt):
                ms = _time_items(st, [it])
                rec = {"n": n, "dtype": dt, "kind": it.kind, "width": len(it.qubits),
                       "nops": len(it.ops), "features": launch_features(it, n, dt).tolist(), "ms": ms}
                if it.kind == "gate":
                    k = G.num_qubits_of(it.U)
                    rec["changed"] = G.changed_amplitudes(it.U, k) << (n - k)
                elif it.kind == "tile":
                    rec["prof"] = list(tile_profile(it, dt))
                recs.append(rec)
            st.close()
    return recs


def dataset(recs: List[Dict]):
    X = np.array([r["features"] for r in recs], dtype=np.float32)
    T = np.array([math.log(r["ms"] / state_gb(r["n"], r["dtype"])) for r in recs], dtype=np.float32)
    return X, T


class CostModel:
    """Trained MLP + the plan chooser that uses it."""

    def __init__(self, params: np.ndarray | None = None):
        self.mlp = MLP(params)

    @classmethod
    def trained(cls, recs: List[Dict], epochs: int = 3000, lr: float = 3e-3, seed: int = 0) -> "CostModel":
        cm = cls(init_params(seed))
        X, T = dataset(recs)
        cm.final_loss = cm.mlp.train(X, T, epochs=epochs, lr=lr)
        return cm

    def predict_ms(self, items: List["P.PlanItem"], n: int, dtype: str) -> np.ndarray:
        if not items:
            return np.zeros(0)
        X = np.stack([launch_features(it, n, dtype) for it in items])
        return np.exp(self.mlp.predict(X).astype(np.float64)) * state_gb(n, dtype)

    def plan_cost(self, plan: List["P.PlanItem"], n: int, dtype: str) -> float:
        return float(self.predict_ms(plan, n, dtype).sum())

    # A candidate replaces the fixed rule only when the model prices it this
    # much cheaper: the held-out error of one launch is a few percent, and
    # the tile-width / low-run variants of one circuit often lie closer than
    # that, where following the model would be a coin toss.
    MARGIN = 0.08

    def choose(self, gates, n: int, dtype: str, candidates: Dict[str, List["P.PlanItem"]] | None = None):
        """Returns (name, plan, predicted ms per candidate)."""
        cands = candidates or candidate_plans(gates, n, dtype)
        costs = {name: self.plan_cost(plan, n, dtype) for name, plan in cands.items()}
        best = min(costs, key=costs.get)
        fixed = fixed_rule_name(dtype)
        if fixed in costs and costs[best] > (1.0 - self.MARGIN) * costs[fixed]:
            best = fixed
        return best, cands[best], costs

    def save(self, path: str, recs: List[Dict] | None = None) -> None:
        with open(path, "w") as f:
            json.dump({"din": DIN, "hidden": HIDDEN, "params": self.mlp.params().tolist(),
                       "records": recs or []}, f)

    _shared: "CostModel | None" = None

    @classmethod
    def shared(cls) -> "CostModel":
        """The model trained on MI355X launch timings that ships with the
        package (loaded once per process)."""
        if cls._shared is None:
            cls._shared = cls.load(DEFAULT_MODEL)
        return cls._shared

    @classmethod
    def load(cls, path: str) -> "CostModel":
        with open(path) as f:
            d = json.load(f)
        assert d["din"] == DIN and d["hidden"] == HIDDEN
        cm = cls(np.array(d["params"], dtype=np.float32))
        cm.records = d.get("records", [])
        return cm


def fixed_rule_name(dtype: str) -> str:
    """The candidate the planner's "tiled" mode runs without a model."""
    return f"tiled_m{P.default_tile_m(dtype)}_l{P.default_low(dtype)}"


def candidate_plans(gates, n: int, dtype: str) -> Dict[str, List["P.PlanItem"]]:
    """plain; fused_k2..6; tiled_m{m}_l{low} for three tile widths and three
    low runs (the qubit permutation into the block: the default run of one
    128 B line and two longer ones); hybrid_m{m}: the default tile passes with
    the gates on more than two qubits as MFMA blocks."""
    cands = {"plain": P.plan_plain(gates, n)}
    for k in (2, 3, 4, 5, 6):
        if k <= n:
            cands[f"fused_k{k}"] = P.plan_fused(gates, n, k)
    if n >= 10:
        m0, l0 = P.default_tile_m(dtype), P.default_low(dtype)
        for m in (m0 - 1, m0, m0 + 1):
            if m > min(n, P.max_tile_m(dtype)):
                continue
            for low in (l0, l0 + 1, l0 + 2):
                cands[f"tiled_m{m}_l{low}"] = P.plan_tiled(gates, n, m, low, merge_1q=True)
        m = min(m0, n)
        cands[f"hybrid_m{m}"] = P.plan_tiled(gates, n, m, l0, merge_1q=True, wide="fused")
    return cands


def timed_plan_ms(state, plan: List["P.PlanItem"], reps: int = 3) -> float:
    """Mean ms of a plan replayed as one graph (after a warm launch); the
    state ends as the plan applied once to |0...0>."""
    prog = P.build_program(state, plan, use_graph=True)
    state.reset_zero()
    prog.run()
    state.sync()
    ms = prog.time(reps)
    state.reset_zero()
    prog.run()
    state.sync()
    prog.close()
    return ms


def evaluate_planners(workload, learned: "CostModel", analytic: "AnalyticModel | None" = None,
                      measure_all: bool = False, log=None) -> Dict:
    """For each (name, n, dtype, gates) circuit: the plan each planner picks
    (learned model, analytic model, fixed rule), the measured time of every
    picked plan (and of every candidate with measure_all), and the picked
    plans' states against the plain state (max |d psi|).  Returns per-circuit
    records and the totals per planner."""
    from .state import HipPureState
    analytic = analytic or AnalyticModel()
    out = {"circuits": [], "totals_ms": {"learned": 0.0, "analytic": 0.0, "fixed_rule": 0.0}}
    if measure_all:
        out["totals_ms"]["best_candidate"] = 0.0
    for name, n, dtype, gates in workload:
        t0 = time.perf_counter()
        cands = candidate_plans(gates, n, dtype)
        t_plan = time.perf_counter() - t0
        pick = {"learned": learned.choose(gates, n, dtype, cands)[0],
                "analytic": analytic.choose(gates, n, dtype, cands)[0],
                "fixed_rule": fixed_rule_name(dtype)}
        ref = HipPureState(n, dtype, fma_tiles=True)
        st = HipPureState(n, dtype, fma_tiles=True)
        P.execute_plan(ref, cands["plain"])
        ref.sync()  # the reference's launches must not overlap the timed plans
        measured, diffs = {}, {}
        keys = list(cands) if measure_all else sorted(set(pick.values()))
        for key in keys:
            measured[key] = timed_plan_ms(st, cands[key])
            if key in pick.values():
                diffs[key] = st.max_abs_diff(ref)
        rec = {"circuit": name, "n": n, "dtype": dtype, "gates": len(gates), "plan_s": t_plan, "pick": pick,
               "measured_ms": {k: round(v, 4) for k, v in measured.items()},
               "max_abs_diff_vs_plain": diffs}
        for who, key in pick.items():
            out["totals_ms"][who] += measured[key]
        if measure_all:
            best = min(measured, key=measured.get)
            rec["best_candidate"] = best
            out["totals_ms"]["best_candidate"] += measured[best]
        out["circuits"].append(rec)
        if log:
            log(f"{name}: learned {pick['learned']} {measured[pick['learned']]:.2f} ms, analytic "
                f"{pick['analytic']} {measured[pick['analytic']]:.2f} ms, fixed {pick['fixed_rule']} "
                f"{measured[pick['fixed_rule']]:.2f} ms")
        st.close()
        ref.close()
    return out


DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "costmodel_mi355x.json")


def main(argv=None) -> int:
    """Retrain the shipped model on this GPU and report the planner
    ablation: python -m qvm_amd.costmodel --out MODEL.json [--report R.json]"""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="where to write the trained model")
    ap.add_argument("--report", default="", help="where to write the ablation on the held-out workload")
    ap.add_argument("--ns", default="16,20,24,28")
    ap.add_argument("--per-config", type=int, default=120)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=21)
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, flush=True)

    t0 = time.perf_counter()
    recs = collect([int(x) for x in args.ns.split(",")], ["f64", "f32"], args.per_config, seed=args.seed)
    log(f"collected {len(recs)} launches in {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(0)
    idx = rng.permutation(len(recs))
    nt = max(1, len(recs) // 5)
    test, train = [recs[i] for i in idx[:nt]], [recs[i] for i in idx[nt:]]
    cm = CostModel.trained(train, epochs=args.epochs)
    Xt, Tt = dataset(test)
    rel = np.exp(np.abs(cm.mlp.predict(Xt) - Tt)) - 1
    am = AnalyticModel()
    arel = np.array([abs(math.log(am.record_seconds(r) * 1e3 / r["ms"])) for r in test])
    log(f"held-out median relative error: MLP {np.median(rel):.3f}, analytic {np.median(np.exp(arel) - 1):.3f}")
    cm = CostModel.trained(recs, epochs=args.epochs)  # the shipped model sees every record
    cm.save(args.out, recs)
    if args.report:
        rep = evaluate_planners(CI.mixed_workload("test"), cm, am, measure_all=True, log=log)
        rep["heldout_median_rel_err"] = {"mlp": float(np.median(rel)), "analytic": float(np.median(np.exp(arel) - 1))}
        rep["records"] = len(recs)
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
        log(f"totals (ms): {rep['totals_ms']}")
    cm.mlp.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
