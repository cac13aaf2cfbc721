"""Learned fusion cost model.

A small MLP predicts the duration of one planned launch (a plain gate, an LDS
tile pass or a fused MFMA block) from features of the launch.  The planner
prices every candidate plan for a circuit -- plain, fused with width
k = 2..6, tiled with m = 10..12 -- by summing the predictions, and runs the
cheapest; the fusion width and the qubits named as tile qubits (the qubit
permutation into LDS) are what the model chooses.  The network is trained and
evaluated by the engine's own HIP kernels (qvm_amd/csrc/mlp.hip) on timings
the engine records with HIP events (``collect``) on states whose tile passes
use the FMA form (the fast mode the planner chooses for).

The target is log(ms per GB of state): launches of any n share one scale, and
the model's prediction for n is exp(y) * state_GB.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from typing import Dict, List, Sequence

import numpy as np

from . import _native as N
from . import circuits as CI
from . import gates as G
from . import planner as P

DIN = 14
HIDDEN = 32


def launch_features(item: "P.PlanItem", n: int, dtype: str) -> np.ndarray:
    """Features of one launch (all in [0, ~1])."""
    f = np.zeros(DIN, dtype=np.float32)
    qs = list(item.qubits)
    kind = item.kind
    f[0] = 1.0
    f[1] = kind == "gate"
    f[2] = kind == "tile"
    f[3] = kind == "fused"
    width = len(qs)
    f[4] = width / 13.0
    f[5] = (len(item.ops) / 128.0) if kind == "tile" else 0.0
    f[6] = min(qs) / 34.0
    f[7] = max(qs) / 34.0
    f[8] = sum(1 for q in qs if q < 4) / 4.0
    f[9] = 1.0 if 0 in qs else 0.0
    if kind == "gate":
        U = item.U
        k = G.num_qubits_of(U)
        f[10] = G.changed_amplitudes(U, k) / float(1 << k)
        f[11] = 1.0 if np.count_nonzero(U - np.diag(np.diag(U))) == 0 else 0.0
    else:
        f[10] = 1.0
        f[11] = 0.0
    f[12] = 1.0 if dtype == "f32" else 0.0
    f[13] = n / 34.0
    return f


def state_gb(n: int, dtype: str) -> float:
    return (1 << n) * (16 if dtype == "f64" else 8) / 1e9


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


def random_launches(n: int, rng: np.random.Generator, count: int) -> List["P.PlanItem"]:
    """Random single launches of every kind the planners emit."""
    items = []
    for _ in range(count):
        r = rng.random()
        if r < 0.35:  # plain gate
            kind = rng.integers(0, 6)
            if kind == 0:
                U, q = G.gate_matrix("RX", [float(rng.random())]), [int(rng.integers(n))]
            elif kind == 1:
                U, q = G.gate_matrix("RZ", [float(rng.random())]), [int(rng.integers(n))]
            elif kind in (2, 3):
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                U, q = G.gate_matrix("CNOT" if kind == 2 else "CZ"), [a, b]
            else:
                k = int(kind - 2)
                Z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
                U, q = np.linalg.qr(Z)[0], [int(x) for x in rng.choice(n, k, replace=False)]
            items.append(P.PlanItem("gate", list(G.nat_tuple(*q)), U=U))
        elif r < 0.7:  # fused block
            k = int(rng.integers(3, 7))
            Z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
            if rng.random() < 0.5:
                qs = [int(x) for x in rng.choice(n, k, replace=False)]
            else:
                lo = int(rng.integers(0, n - k + 1))
                qs = list(range(lo, lo + k))
            items.append(P.PlanItem("fused", qs, U=np.linalg.qr(Z)[0]))
        else:  # tile pass over a random layered fragment
            m = int(rng.choice([10, 11, 12]))
            depth = int(rng.integers(1, 6))
            gates = CI.layered(m, depth, seed=int(rng.integers(1 << 30)))
            hi = sorted(int(x) for x in rng.choice(np.arange(4, n), m - 4, replace=False))
            tq = list(range(4)) + hi
            relabel = {j: tq[j] for j in range(m)}
            g2 = [(U, [relabel[x] for x in q]) for U, q in gates]
            keep = int(rng.integers(4, len(g2) + 1))
            plan = P.plan_tiled(g2[:keep], n, m, 4, reorder=False)
            items.extend(it for it in plan if it.kind == "tile")
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
            for it in random_launches(n, rng, per_config):
                ms = _time_items(st, [it])
                recs.append({"n": n, "dtype": dt, "kind": it.kind, "width": len(it.qubits),
                             "nops": len(it.ops), "features": launch_features(it, n, dt).tolist(), "ms": ms})
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

    def choose(self, gates, n: int, dtype: str, candidates: Dict[str, List["P.PlanItem"]] | None = None):
        """Returns (name, plan, predicted ms per candidate)."""
        cands = candidates or candidate_plans(gates, n, dtype)
        costs = {name: self.plan_cost(plan, n, dtype) for name, plan in cands.items()}
        best = min(costs, key=costs.get)
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


def candidate_plans(gates, n: int, dtype: str) -> Dict[str, List["P.PlanItem"]]:
    low = P.default_low(dtype)
    cands = {"plain": P.plan_plain(gates, n)}
    for k in (2, 3, 4, 5, 6):
        if k <= n:
            cands[f"fused_k{k}"] = P.plan_fused(gates, n, k)
    if n >= 10:
        for m in (10, 11, 12) if dtype == "f64" else (11, 12, 13):
            if m <= n:
                cands[f"tiled_m{m}"] = P.plan_tiled(gates, n, m, low)
    return cands


DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "costmodel_mi355x.json")
