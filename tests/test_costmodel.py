"""The learned fusion cost model: its CPU oracle, its GPU kernels, and what
the trained model gains over the planner without it."""

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from qvm_amd import circuits as CI
from qvm_amd import costmodel as CM
from qvm_amd import planner as P


def _data(nb=48, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.random((nb, CM.DIN)).astype(np.float32)
    T = (np.sin(3 * X[:, 1]) + X[:, 4] ** 2 - 0.5 * X[:, 7]).astype(np.float32)
    return X, T


def test_mlp_oracle_gradient_matches_finite_differences():
    din, h = 5, 7
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.5, MO.num_params(din, h))
    X = rng.normal(size=(9, din))
    T = rng.normal(size=9)
    loss, g = MO.loss_and_grad(p, X, T, din, h)
    eps = 1e-6
    for i in rng.choice(p.size, 40, replace=False):
        q = p.copy()
        q[i] += eps
        lp, _ = MO.loss_and_grad(q, X, T, din, h)
        q[i] -= 2 * eps
        lm, _ = MO.loss_and_grad(q, X, T, din, h)
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-7 * max(1.0, abs(g[i]))
    assert abs(loss - 0.5 * np.mean((MO.forward(p, X, din, h) - T) ** 2)) < 1e-12


def test_launch_features_and_candidates():
    n = 14
    gates = CI.random_mix(n, 4, seed=2)
    cands = CM.candidate_plans(gates, n, "f64")
    assert {"plain", "fused_k4", CM.fixed_rule_name("f64"), "tiled_m10_l3", "tiled_m12_l5", "hybrid_m11"} <= set(cands)
    for name, plan in cands.items():
        for it in plan:
            f = CM.launch_features(it, n, "f64")
            assert f.shape == (CM.DIN,) and np.all(np.isfinite(f)) and f.max() <= 2.0 + 1e-6
    assert len(cands["fused_k5"]) < len(cands["plain"])
    # the low run is the qubit permutation into the block: every pass of
    # tiled_m*_l{L} keeps qubits 0..L-1
    for L in (3, 4, 5):
        for it in cands[f"tiled_m11_l{L}"]:
            if it.kind == "tile":
                assert CM.low_run(it.qubits) >= L


def test_hybrid_plan_runs_wide_gates_on_the_matrix_cores():
    n = 14
    gates = CI.dense_blocks(n, 2, 5, seed=3)
    hyb = CM.candidate_plans(gates, n, "f64")["hybrid_m11"]
    wide = [it for it in hyb if it.kind != "tile"]
    assert wide and all(it.kind == "fused" and len(it.qubits) == 5 for it in wide)
    # the analytic model prices from first principles: the fused k=6 block
    # is MFMA-bound at FP64, a 1q gate is HBM-bound
    am = CM.AnalyticModel()
    t1 = am.seconds("gate", 30, "f64", 1, changed=1 << 30)
    t6 = am.seconds("fused", 30, "f64", 6)
    assert abs(t1 - (am.OVERHEAD_S + 2 * 16 * 2 ** 30 / am.STREAM_BPS)) < 1e-12
    assert t6 > am.seconds("fused", 30, "f64", 4) >= t1


@pytest.mark.gpu
def test_mlp_gpu_forward_and_gradient_match_oracle():
    X, T = _data()
    p0 = CM.init_params(3)
    m = CM.MLP(p0)
    y = m.predict(X)
    y_ref = MO.forward(p0.astype(np.float64), X, CM.DIN, CM.HIDDEN)
    assert np.max(np.abs(y - y_ref)) < 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = m.gradient(X, T)
    loss_ref, g_ref = MO.loss_and_grad(p0.astype(np.float64), X, T, CM.DIN, CM.HIDDEN)
    assert abs(loss - loss_ref) < 1e-4 * max(1.0, loss_ref)
    assert np.max(np.abs(g - g_ref)) < 2e-4 * max(1.0, np.max(np.abs(g_ref)))
    m.close()


@pytest.mark.gpu
def test_mlp_training_fits():
    X, T = _data(96, seed=4)
    m = CM.MLP(CM.init_params(0))
    l0, _ = m.gradient(X, T)
    l1 = m.train(X, T, epochs=1500, lr=3e-3)
    assert l1 < 0.1 * l0
    p = m.params()
    l_ref, _ = MO.loss_and_grad(p.astype(np.float64), X, T, CM.DIN, CM.HIDDEN)
    l_now, _ = m.gradient(X, T)
    assert abs(l_now - l_ref) < 1e-3 * max(1.0, l_ref)
    m.close()


@pytest.mark.gpu
def test_cost_model_trains_on_engine_timings():
    """Timings the engine records train the network on the GPU to a useful
    accuracy (in-sample median relative error below 50 %)."""
    recs = CM.collect([20], ["f64", "f32"], per_config=60, seed=1)
    cm = CM.CostModel.trained(recs, epochs=2500)
    X, T = CM.dataset(recs)
    rel = np.exp(np.abs(cm.mlp.predict(X) - T)) - 1
    print(f"cost model: {len(recs)} launches, median relative error {np.median(rel):.2f}", flush=True)
    assert np.median(rel) < 0.5
    cm.mlp.close()


@pytest.mark.gpu
def test_learned_planner_ablation():
    """On a held-out workload whose best plan differs by circuit (random
    1q/2q mixes, dense 5- and 6-qubit blocks, ZZ-phase heavy circuits, small
    states), every candidate is timed.  The shipped model's picks must run
    faster in total than the fixed rule's (tile passes at the default width)
    and no slower than the analytic model's (3 % timing slack), must take at
    least half of the gain the best candidates offer over the fixed rule, and
    every picked plan's state must equal the plain state."""
    cm = CM.CostModel.shared()
    rep = CM.evaluate_planners(CI.mixed_workload("test", seed=1900), cm, CM.AnalyticModel(), measure_all=True,
                               log=lambda *a: print(*a, flush=True))
    tot = rep["totals_ms"]
    print(f"totals (ms): {tot}", flush=True)
    for rec in rep["circuits"]:
        tol = 1e-11 if rec["dtype"] == "f64" else 1e-6
        for key, d in rec["max_abs_diff_vs_plain"].items():
            assert d < tol, (rec["circuit"], key, d)
    assert tot["learned"] < tot["fixed_rule"]
    assert tot["learned"] <= 1.03 * tot["analytic"]
    assert tot["fixed_rule"] - tot["learned"] >= 0.5 * (tot["fixed_rule"] - tot["best_candidate"])
