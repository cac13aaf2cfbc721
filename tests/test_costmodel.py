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
    assert {"plain", "fused_k4", "tiled_m10"} <= set(cands)
    for name, plan in cands.items():
        for it in plan:
            f = CM.launch_features(it, n, "f64")
            assert f.shape == (CM.DIN,) and np.all(np.isfinite(f)) and f.max() <= 1.0 + 1e-6
    assert len(cands["fused_k5"]) < len(cands["plain"])


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
def test_learned_planner_beats_fixed_width():
    """Train on engine-recorded timings, then on held-out circuits compare the
    measured time of the model's choice with the fixed-width (k=4) fusion
    planner and with plain gate-by-gate application."""
    from qvm_amd.state import HipPureState
    n = 22
    recs = CM.collect([n], ["f64", "f32"], per_config=70, seed=1)
    cm = CM.CostModel.trained(recs, epochs=2500)
    X, T = CM.dataset(recs)
    pred = cm.mlp.predict(X)
    rel = np.exp(np.abs(pred - T)) - 1
    print(f"cost model: {len(recs)} launches, median relative error {np.median(rel):.2f}", flush=True)
    assert np.median(rel) < 0.5
    totals = {"learned": 0.0, "fixed_k4": 0.0, "plain": 0.0}
    for i, dtype in enumerate(["f64", "f32", "f64"]):
        gates = CI.random_mix(n, 12, seed=100 + i)
        st = HipPureState(n, dtype, fma_tiles=True)
        cands = CM.candidate_plans(gates, n, dtype)
        name, plan, costs = cm.choose(gates, n, dtype, cands)
        measured = {}
        for key in set([name, "fused_k4", "plain"]):
            prog = P.build_program(st, cands[key], use_graph=True)
            prog.run()
            st.sync()
            measured[key] = prog.time(3)
            prog.close()
        totals["learned"] += measured[name]
        totals["fixed_k4"] += measured["fused_k4"]
        totals["plain"] += measured["plain"]
        print(f"circuit {i} {dtype}: learned picks {name} ({measured[name]:.3f} ms), "
              f"fixed k=4 {measured['fused_k4']:.3f} ms, plain {measured['plain']:.3f} ms", flush=True)
        st.close()
    print(f"totals: {totals}", flush=True)
    assert totals["learned"] <= 1.05 * min(totals["fixed_k4"], totals["plain"])
    cm.mlp.close()
