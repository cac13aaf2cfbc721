#!/usr/bin/env python3
"""Benchmark of the pure-state path on one MI355X.

Workload (one step): a 30-qubit FP64 statevector (16 GiB, resident in HBM)
  (a) layered circuit, 40 layers of RX(a) q, RZ(b) q on every qubit plus
      brickwork CNOTs (2400 1-qubit gates + 580 CNOTs), from |0...0>;
  (b) QFT-30 (30 H, 435 CPHASE, 15 SWAP) from |0...0>;
  then 10^6 shots sampled from the FP64 prefix sum of (b)'s state.
Gates run on the plain path: one engine launch per gate, the arithmetic of
the reference's serial kernels.

value = amp-updates/s of the step: sum over gates of changed(g) (2^n for
dense/diagonal 1q gates, 2^(n-1) for CNOT/SWAP, 2^(n-2) for CPHASE) divided
by the step's wall time.  %HBM = 2 * changed * 16 B / time / 8 TB/s.

Also reported (measured once each, outside the timed loop):
  parity           -- the timed shots against the closed form of QFT|0>'s
                      uniform CDF (ceil(u 2^n) - 1), near-boundary counts.
  paths.tiled_f64  -- the same circuits as register-tile passes (bit-identical),
  paths.fused_f64  -- dense 4-qubit blocks on MFMA,
  paths.fused_f32 / tiled_f32 -- FP32 states, with infidelity, max |d psi|
                      and total-variation distance against the FP64 result.
  precision        -- Haar-U3 + CZ-matching circuits, depth 20: FP32 against
                      FP64 at n = 30, FP32 alone at n = 34 (128 GiB).
  batch            -- 512 QAOA MaxCut instances (<H_C> + 1024 shots each) in
                      one graph launch, FP64 and FP32, spot-checked against
                      single-state plain runs of three instances.
  learned          -- the shipped MLP cost model's plan for depth-60 random
                      circuits at n = 30 against fixed-rule plans and the
                      plain FP64 state.
  cpu_baseline     -- the oracle (a C restatement of the reference CPU path)
                      on the box's host cores: one gate of each kind in the
                      step, timed and extrapolated to the step's gate mix, in
                      compiled and interp-avx mode on all usable cores and in
                      compiled mode on one thread.
  config1          -- the 20-qubit CPU-runnable configuration in full on
                      both sides.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK = 8.0e12  # B/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--shots", type=int, default=1_000_000)
    ap.add_argument("--no-paths", action="store_true", help="skip the tiled/fused/FP32 comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU time bound of the standalone workloads")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle worker threads (0: the logical cores sysconf reports, capped at the "
                         "cores the process's cgroup quota grants)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the precision / batch / learned keys of the headline line")
    ap.add_argument("--workload", default="circuits", choices=["circuits", "precision", "batch", "learned"],
                    help="circuits: the headline line (default); precision: FP32 vs FP64 report; "
                         "batch: 512 QAOA instances in one graph")
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--learn-samples", type=int, default=120, help="timed launches per (n, dtype) for the cost model")
    ap.add_argument("--save-model", default="",
                    help="--workload learned: retrain the cost model first and write it (params + records) here")
    ap.add_argument("--big-n", type=int, default=34, help="FP32-only size of the precision report")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path "
              "(sharding over the top qubits is a design, not built); run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "precision":
        return precision(args)
    if args.workload == "batch":
        return batch(args)
    if args.workload == "learned":
        return learned(args)

    import numpy as np

    from qvm_amd import circuits as C
    from qvm_amd import gates as G
    from qvm_amd import planner as P
    from qvm_amd.state import HipPureState

    n = args.n
    ga = C.layered(n, args.layers, seed=1)
    gb = C.qft(n)
    upd_a, upd_b = C.amp_updates(ga, n), C.amp_updates(gb, n)
    upd = upd_a + upd_b
    u = C.uniforms(20250, args.shots)
    nat_a = [(U, list(G.nat_tuple(*q))) for U, q in ga]
    nat_b = [(U, list(G.nat_tuple(*q))) for U, q in gb]

    st = HipPureState(n, "f64")

    def run_plain(state, gates):
        state.reset_zero()
        for U, q in gates:
            state.apply_matrix(U, q)

    def step():
        run_plain(st, nat_a)
        run_plain(st, nat_b)
        return st.sample(u)

    log(f"bench: n={n} layered gates={len(ga)} qft gates={len(gb)} amp-updates/step={upd:.4e}")
    for w in range(args.warmup):
        step()
        log(f"warmup {w} done")
    st.sync()
    st.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        shots = step()
        log(f"step {k} done")
    st.sync()
    t1 = time.perf_counter()
    prof = st.profile_summary()
    st.profile(False)
    step_s = (t1 - t0) / args.steps
    value = upd / step_s
    bytes_alg = 2 * upd * 16

    # parity of the timed result: QFT|0> is the uniform superposition, so
    # the draws' closed form is ceil(u 2^n) - 1 (rule cdf >= u); the same
    # draws again with near-boundary flags must give the timed indices
    amps = st.read_amplitudes(0, 4096)
    qft_err = float(np.max(np.abs(amps - 2.0 ** (-n / 2))))
    norm_b = st.norm()
    again, near = st.sample_ex(u, "ge", 1e-15)
    closed = np.clip(np.ceil(u * float(1 << n)) - 1, 0, (1 << n) - 1).astype(np.int64)
    dif = shots.astype(np.int64) - closed
    bad = np.nonzero(dif)[0]
    grid = np.abs(u[bad] * float(1 << n) - np.round(u[bad] * float(1 << n))) / float(1 << n)
    shot_parity = {
        "rule": "min{i : cdf[i] >= u} (src/measurement.lisp:246-288)",
        "reference": "closed form ceil(u 2^n) - 1 of the uniform CDF (i + 1) / 2^n",
        "shots": int(u.size),
        "shots_mismatch": int(bad.size),
        "max_abs_index_diff": int(np.abs(dif).max()) if u.size else 0,
        "mismatch_max_distance_to_grid": float(grid.max()) if bad.size else 0.0,
        "near_boundary_1e-15": int(near.sum()),
        "timed_shots_equal_flagged_rerun": bool(np.array_equal(again, shots)),
    }

    # dominant kernel: family with the largest total time
    dom, (dcount, dms) = max(prof.items(), key=lambda kv: kv[1][1])
    per_launch_ms = dms / dcount
    unit_bytes = 2 * (1 << n) * 16  # dense 1q / diagonal sweep: read + write every amplitude
    roofline = {
        "kernel": {"dense1": "1-qubit dense family: qvmhip::k_dense1_lane<double> (target < 6), "
                             "qvmhip::k_dense1_xchg<double> (>= 6)",
                   "diag_full": "qvmhip::k_diag_full<double, true, 1>"}.get(dom, dom),
        "bound": "hbm",
        "achieved": round(unit_bytes / (per_launch_ms * 1e-3) / 1e9, 1),
        "peak": 8000.0,
        "unit": "GB/s",
        "frac": round(unit_bytes / (per_launch_ms * 1e-3) / HBM_PEAK, 4),
        "traffic": None,
        "launches": dcount,
        "avg_launch_ms": round(per_launch_ms, 4),
        "algorithmic_bytes_per_launch": unit_bytes,
    }
    pmc = os.path.join(REPO, "profiles", f"pmc_{dom}_n30.json")
    if os.path.exists(pmc) and n == 30:
        # HBM bytes per launch from the committed rocprofv3 --pmc summary of
        # the same kernels (FETCH_SIZE / WRITE_SIZE scaled by a copy of known
        # size); counters are not collected inside this timed run
        with open(pmc) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")
        roofline["traffic_source"] = f"profiles/pmc_{dom}_n30.json (rocprofv3 --pmc)"
    kernels = {k: {"launches": c, "total_ms": round(t, 2), "avg_ms": round(t / c, 4)} for k, (c, t) in prof.items()}

    out = {
        "metric": "amp-updates/s + %HBM peak, 30q FP64 plain vs fused/FP32 at infidelity≤1e-6",
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded (splitmix64) layered RX/RZ+CNOT and QFT circuits, seeded uniform draws",
        "config": {
            "workload": f"{n}-qubit FP64 statevector: layered RX/RZ+CNOT depth {args.layers} + QFT-{n}, "
                        f"then {args.shots} shots from the FP64 prefix sum; plain gate-by-gate path",
            "n_qubits": n, "layers": args.layers, "shots": args.shots,
            "gates": len(ga) + len(gb), "amp_updates_per_step": upd,
        },
        "hbm_frac": round(bytes_alg / step_s / HBM_PEAK, 4),
        "roofline": roofline,
        "kernels": kernels,
        "parity": {"qft_uniform_max_abs_err": qft_err, "norm": norm_b, "shots_vs_closed_form": shot_parity},
    }

    out["ceilings"] = ceilings(st, prof, n, HipPureState)

    t_sec = {}
    if not args.no_paths:
        t0 = time.perf_counter()
        out["paths"] = run_paths(args, n, ga, gb, nat_a, upd_a, upd_b, u, st, HipPureState, P, np)
        t_sec["paths"] = time.perf_counter() - t0
    st.close()
    if not args.no_extras and n == 30:
        for key, fn in (("precision", precision_key), ("batch", batch_key), ("learned", learned_key)):
            t0 = time.perf_counter()
            out[key] = fn(args)
            t_sec[key] = time.perf_counter() - t0
            log(f"{key} done in {t_sec[key]:.1f} s")
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        threads, cores_info = cpu_threads(args.cpu_threads)
        out["cpu_baseline"] = cpu_baseline(n, ga, gb, threads, cores_info)
        out["config1"] = config1(threads, HipPureState)
        t_sec["cpu_baseline"] = time.perf_counter() - t0
    out["section_wall_s"] = {k: round(v, 1) for k, v in t_sec.items()}
    print(json.dumps(out))
    return 0


def ceilings(st, prof, n, HipPureState):
    """What the part streams at, measured in this run: a device-to-device
    copy of the whole state (hipMemcpyAsync, read + write), and the in-place
    sweep of the full-diagonal gate kernel (RZ: one read and one write per
    amplitude, four multiplies), from the timed steps' own launches; plus the
    device as rocminfo names it."""
    import subprocess
    res = {}
    other = HipPureState(n, "f64")
    other.copy_from(st)
    other.sync()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        other.copy_from(st)
    other.sync()
    dt = (time.perf_counter() - t0) / reps
    other.close()
    nbytes = 2 * (1 << n) * 16
    res["d2d_copy_GBps"] = round(nbytes / dt / 1e9, 1)
    if "diag_full" in prof:
        c, t = prof["diag_full"]
        res["inplace_sweep_GBps"] = round(nbytes / (t / c * 1e-3) / 1e9, 1)
        res["inplace_sweep_kernel"] = "qvmhip::k_diag_full<double, true, 1> (RZ)"
    try:
        txt = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
        names = [ln.split(":", 1)[1].strip() for ln in txt.splitlines() if "Marketing Name" in ln]
        cus = [ln.split(":", 1)[1].strip() for ln in txt.splitlines() if "Compute Unit" in ln]
        gfx = [ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ln.strip().startswith("Name:") and "gfx" in ln]
        res["rocminfo"] = {"marketing_names": names, "compute_units": cus, "gfx": gfx}
    except (OSError, subprocess.SubprocessError) as e:
        res["rocminfo"] = f"unavailable: {e}"
    return res


def run_paths(args, n, ga, gb, nat_a, upd_a, upd_b, u, st, HipPureState, P, np):
    """Tiled / fused / FP32 forms of the same circuits against the plain FP64 state."""
    res = {}
    # plain FP64 reference results for (a) and (b), kept on device
    ref_a = HipPureState(n, "f64")
    for U, q in nat_a:
        ref_a.apply_matrix(U, q)
    ref_a.sync()
    st.reset_zero()  # st now holds the plain QFT result? rebuild it explicitly
    for U, q in gb:
        st.apply_matrix(U, list(reversed(q)))
    st.sync()
    ref_b = st
    p64 = ref_a.sample(u)

    def timed_program(state, plan):
        prog = P.build_program(state, plan, use_graph=True)
        state.reset_zero()
        prog.run()  # warm
        state.sync()
        state.reset_zero()
        state.sync()
        t0 = time.perf_counter()
        prog.run()
        state.sync()
        dt = time.perf_counter() - t0
        prog.close()
        return dt

    def record(name, state, plan_a, plan_b):
        ta = timed_program(state, plan_a)
        da = state.max_abs_diff(ref_a)
        nrm = state.norm()
        ip = state.inner_product(ref_a) / nrm  # fidelity of the normalised state
        if state.dtype == "f32":
            state.normalize(nrm)  # FP64-accumulated renormalisation before sampling
        pa = state.sample(u)
        tv = state.tv_distance(ref_a)  # TV of the outcome distribution |psi|^2
        tb = timed_program(state, plan_b)
        db = state.max_abs_diff(ref_b)
        tot = ta + tb
        r = {"amp_updates_per_s": (upd_a + upd_b) / tot, "s_layered": ta, "s_qft": tb,
             "sweeps_layered": len(plan_a), "sweeps_qft": len(plan_b),
             "max_abs_diff_vs_plain_f64": max(da, db),
             "infidelity_vs_plain_f64": max(0.0, 1.0 - abs(ip) ** 2),
             "shots_differing_vs_f64": int(np.count_nonzero(pa != p64))}
        r["tv_distance_vs_plain_f64"] = tv
        if state.dtype == "f32":
            r["norm_before_renormalisation"] = nrm
            # same uniforms: shots land on the same or a neighbouring index
            r["shots_median_index_shift"] = float(np.median(np.abs(pa.astype(np.int64) - p64.astype(np.int64))))
        log(f"path {name}: {r['amp_updates_per_s']:.3e} amp-updates/s, diff {r['max_abs_diff_vs_plain_f64']:.2e}")
        return r

    s2 = HipPureState(n, "f64")
    m64, l64 = P.default_tile_m("f64"), P.default_low("f64")
    res["tiled_exact_f64"] = record("tiled_exact_f64", s2, P.plan_tiled(ga, n, m64, l64, reorder=False),
                                    P.plan_tiled(gb, n, m64, l64, reorder=False))
    res["tiled_exact_f64"]["bit_identical_to_plain"] = res["tiled_exact_f64"]["max_abs_diff_vs_plain_f64"] == 0.0
    res["fused_f64"] = record("fused_f64", s2, P.plan_fused(ga, n, 4), P.plan_fused(gb, n, 4))
    s2.close()
    s2 = HipPureState(n, "f64", fma_tiles=True)  # reordered passes, FMA-contracted
    res["tiled_f64"] = record("tiled_f64", s2, P.plan_tiled(ga, n, m64, l64, merge_1q=True),
                              P.plan_tiled(gb, n, m64, l64, merge_1q=True))
    s2.close()
    s3 = HipPureState(n, "f32", fma_tiles=True)
    m32 = P.default_tile_m("f32")
    res["tiled_f32"] = record("tiled_f32", s3, P.plan_tiled(ga, n, m32, P.default_low("f32"), merge_1q=True),
                              P.plan_tiled(gb, n, m32, P.default_low("f32"), merge_1q=True))
    res["fused_f32"] = record("fused_f32", s3, P.plan_fused(ga, n, 5), P.plan_fused(gb, n, 5))
    s3.close()
    ref_a.close()
    return res


def _timed_plan(state, plan, P, warm=True):
    """One timed graph launch of a plan from |0...0> (after a warm one)."""
    prog = P.build_program(state, plan, use_graph=True)
    if warm:
        state.reset_zero()
        prog.run()
        state.sync()
    state.reset_zero()
    state.sync()
    t0 = time.perf_counter()
    prog.run()
    state.sync()
    dt = time.perf_counter() - t0
    prog.close()
    return dt


def _vs(state, ref, nrm=None):
    """Infidelity, max |d psi| and TV distance of state against ref (the
    normalised state for the infidelity and the TV distance)."""
    nrm = state.norm() if nrm is None else nrm
    ip = state.inner_product(ref) / (nrm * ref.norm())
    d = state.max_abs_diff(ref)
    if state.dtype == "f32":
        state.normalize(nrm)
    return {"infidelity": max(0.0, 1.0 - abs(ip) ** 2), "max_abs_amp_err": d, "tv_distance": state.tv_distance(ref),
            "norm_before_renormalisation": nrm}


def precision_key(args):
    """Config 3 in the headline line: Haar-U3 + CZ matching, depth 20.  n=30:
    FP64 plain (the reference state), FP32 tiled against it; n=34: FP32 tiled
    (128 GiB), its norm and a closed-form-free check (the run and its inverse
    pass return to |0...0>)."""
    from qvm_amd import circuits as C
    from qvm_amd import gates as G
    from qvm_amd import planner as P
    from qvm_amd.state import HipPureState
    n = 30
    gates = C.haar_cz(n, 20, seed=3)
    upd = C.amp_updates(gates, n)
    res = {"circuit": "Haar U3 on every qubit + CZ on a seeded perfect matching, depth 20",
           "n30_gates": len(gates), "n30_amp_updates": upd}
    ref = HipPureState(n, "f64")
    ref.sync()
    t0 = time.perf_counter()
    for U, q in gates:
        ref.apply_matrix(U, G.nat_tuple(*q))
    ref.sync()
    t64 = time.perf_counter() - t0
    res["f64_plain_n30"] = {"s": t64, "amp_updates_per_s": upd / t64}
    s32 = HipPureState(n, "f32", fma_tiles=True)
    plan = P.plan_tiled(gates, n, P.default_tile_m("f32"), P.default_low("f32"), merge_1q=True)
    t32 = _timed_plan(s32, plan, P)
    res["f32_tiled_n30"] = {"s": t32, "amp_updates_per_s": upd / t32, "passes": len(plan),
                            "speedup_vs_f64_plain": t64 / t32, **_vs(s32, ref)}
    s32.close()
    ref.close()
    log(f"precision n=30: f64 plain {t64:.2f} s, f32 tiled {t32:.3f} s, "
        f"infidelity {res['f32_tiled_n30']['infidelity']:.2e}")
    nb = args.big_n
    if nb:
        gb = C.haar_cz(nb, 20, seed=3)
        updb = C.amp_updates(gb, nb)
        big = HipPureState(nb, "f32", fma_tiles=True)
        plan = P.plan_tiled(gb, nb, P.default_tile_m("f32"), P.default_low("f32"), merge_1q=True)
        tb = _timed_plan(big, plan, P, warm=False)  # the same kernels ran warm at n = 30
        nrm = big.norm()
        inv = [(G.dagger(U), q) for U, q in reversed(gb)]
        prog = P.build_program(big, P.plan_tiled(inv, nb, P.default_tile_m("f32"), P.default_low("f32"),
                                                 merge_1q=True), use_graph=True)
        prog.run()
        big.sync()
        prog.close()
        a0 = complex(big.read_amplitudes(0, 1)[0])
        res[f"f32_tiled_n{nb}"] = {"s": tb, "amp_updates_per_s": updb / tb, "gates": len(gb), "passes": len(plan),
                                   "state_GiB": (1 << nb) * 8 / 2 ** 30, "norm": nrm,
                                   "roundtrip_abs_amp0": abs(a0),
                                   "roundtrip_total_probability": big.total_probability()}
        log(f"precision n={nb}: f32 tiled {tb:.2f} s, norm {nrm:.9f}, round trip |a0| {abs(a0):.9f}")
        big.close()
    return res


def batch_key(args):
    """Config 4 in the headline line: 512 QAOA MaxCut instances, <H_C> +
    1024 shots each, FP64 and FP32, one hipGraph launch per run; the three
    smallest instances are re-run as single states on the plain path (the
    engine's bit-exact gate kernels) and compared: amplitudes, <H_C> and the
    same draws (near-boundary counts from the single-state sampler)."""
    import numpy as np
    from qvm_amd import circuits as C
    from qvm_amd import gates as G
    from qvm_amd.batch import qaoa_batch
    from qvm_amd.state import HipPureState
    inst = C.qaoa_instances(args.instances, seed=4)
    circuits = [C.qaoa_maxcut(n, e, g, b) for n, e, g, b in inst]
    upd = sum(C.amp_updates(c, n) for c, (n, _, _, _) in zip(circuits, inst))
    res = {"instances": len(inst), "amp_updates_per_run": upd, "shots_per_instance": 1024}
    check = sorted(range(len(inst)), key=lambda i: inst[i][0])[:3]
    hcs = {}
    for dtype in ("f64", "f32"):
        t0 = time.perf_counter()
        b, _, u = qaoa_batch(inst, dtype, shots=1024, seed=9, use_graph=True, fma_tiles=True, tile_m="auto")
        t_build = time.perf_counter() - t0
        b.run()
        b.sync()
        t0 = time.perf_counter()
        b.run()
        b.sync()
        dt = time.perf_counter() - t0
        hc, shots = b.results()
        hcs[dtype] = hc
        r = {"s_per_run": dt, "instances_per_s": len(inst) / dt, "amp_updates_per_s": upd / dt,
             "tile_m": b.tile_m, "autotune_s_per_run_on_sample": {str(k): v for k, v in b.tile_timings.items()},
             "plan_autotune_capture_s": t_build}
        if dtype == "f64":
            spot = []
            for i in check:
                n, edges, _, _ = inst[i]
                st = HipPureState(n, "f64")
                for U, q in circuits[i]:
                    st.apply_matrix(U, G.nat_tuple(*q))
                ui = u[i * 1024:(i + 1) * 1024]
                got, near = st.sample_ex(ui, "ge", 1e-15)
                spot.append({"instance": i, "n": n, "max_abs_amp_diff": float(np.max(np.abs(b.amplitudes(i) - st.amplitudes()))),
                             "hc_abs_diff": abs(float(hc[i]) - st.expectation_zz(edges)),
                             "shots_mismatch": int(np.count_nonzero(got != shots[i])),
                             "near_boundary_1e-15": int(near.sum())})
                st.close()
            r["spot_check_vs_single_state_plain"] = spot
        res[dtype] = r
        log(f"batch {dtype}: {dt * 1e3:.1f} ms/run ({len(inst) / dt:.0f} instances/s), tile width {b.tile_m}")
        b.close()
    res["f32_max_abs_hc_diff_vs_f64"] = float(np.max(np.abs(hcs["f32"] - hcs["f64"])))
    return res


def learned_key(args):
    """Config 5 in the headline line: the shipped cost model (an MLP trained
    on launch timings the engine recorded, through the engine's own HIP
    kernels) picks, for each circuit of a workload whose best plan differs by
    circuit (depth-20 random 60 % Haar U3 / 40 % CNOT-or-CZ circuits at n = 30
    in FP64 and FP32, dense 6-qubit blocks at n = 28, a ZZ-phase heavy circuit
    at n = 30, a small state), among plain, fused k = 2..6, tiled (three
    widths x three qubit permutations) and hybrid plans.  Its picks are timed
    against the fixed rule (tile passes at the default width) and the
    analytic bytes + issue model, and every picked plan's state is compared
    with the plain state."""
    from qvm_amd import circuits as C
    from qvm_amd import costmodel as CM
    rep = CM.evaluate_planners(C.mixed_workload("bench"), CM.CostModel.shared(), CM.AnalyticModel(), log=log)
    tot = rep["totals_ms"]
    rep["speedup_vs_fixed_rule"] = tot["fixed_rule"] / tot["learned"]
    rep["speedup_vs_analytic"] = tot["analytic"] / tot["learned"]
    return rep


def precision(args) -> int:
    """FP32 against FP64 on Haar-U3 + CZ-matching circuits of depth 20: n=30
    in both dtypes (infidelity, max |d psi|, TV distance of the outcome
    distributions), and the FP32-only size (default n=34, 128 GiB) timed on
    the plain and tiled paths."""
    from qvm_amd import circuits as C
    from qvm_amd import gates as G
    from qvm_amd import planner as P
    from qvm_amd.state import HipPureState

    def run(state, gates, mode):
        state.reset_zero()
        state.sync()
        t0 = time.perf_counter()
        if mode == "plain":
            for U, q in gates:
                state.apply_matrix(U, G.nat_tuple(*q))
        else:
            m = P.default_tile_m(state.dtype)
            prog = P.build_program(state, P.plan_tiled(gates, state.num_qubits, m, P.default_low(state.dtype)),
                                   use_graph=True)
            state.reset_zero()
            state.sync()
            t0 = time.perf_counter()
            prog.run()
        state.sync()
        dt = time.perf_counter() - t0
        if mode != "plain":
            prog.close()
        return dt

    n = args.n
    gates = C.haar_cz(n, 20, seed=3)
    upd = C.amp_updates(gates, n)
    out = {"metric": "FP32 vs FP64 precision, Haar-U3 + CZ matching, depth 20", "n": n,
           "gates": len(gates), "amp_updates": upd, "results": {}}
    ref = HipPureState(n, "f64")
    t64 = run(ref, gates, "plain")
    log(f"precision: f64 plain {t64:.2f} s")
    out["results"]["f64_plain"] = {"s": t64, "amp_updates_per_s": upd / t64, "norm": ref.norm()}
    for mode in ("plain", "tiled"):
        s32 = HipPureState(n, "f32", fma_tiles=True)  # the FMA form of the tile passes
        dt = run(s32, gates, mode)
        nrm = s32.norm()
        ip = s32.inner_product(ref) / nrm
        s32.normalize(nrm)
        r = {"s": dt, "amp_updates_per_s": upd / dt, "norm_before_renormalisation": nrm,
             "infidelity_vs_f64": max(0.0, 1.0 - abs(ip) ** 2),
             "max_abs_amp_err_vs_f64": s32.max_abs_diff(ref),
             "tv_distance_vs_f64": s32.tv_distance(ref)}
        log(f"precision: f32 {mode} {dt:.2f} s infidelity {r['infidelity_vs_f64']:.2e}")
        out["results"][f"f32_{mode}"] = r
        s32.close()
    t64t = run(ref, gates, "tiled")
    out["results"]["f64_tiled"] = {"s": t64t, "amp_updates_per_s": upd / t64t}
    ref.close()
    nb = args.big_n
    if nb:
        gb = C.haar_cz(nb, 20, seed=3)
        updb = C.amp_updates(gb, nb)
        big = HipPureState(nb, "f32", fma_tiles=True)
        for mode in ("tiled", "plain"):
            dt = run(big, gb, mode)
            nrm = big.norm()
            out["results"][f"f32_n{nb}_{mode}"] = {"s": dt, "amp_updates_per_s": updb / dt, "gates": len(gb),
                                                   "norm": nrm, "state_GiB": (1 << nb) * 8 / 2 ** 30}
            log(f"precision: f32 n={nb} {mode} {dt:.2f} s norm {nrm:.9f}")
        big.close()
    print(json.dumps(out))
    return 0


def batch(args) -> int:
    """Config 4: QAOA MaxCut instances (n even in [10, 24], random 3-regular
    graphs, p in 1..4), <H_C> + 1024 shots each, FP64 and FP32, every run one
    hipGraph launch."""
    import numpy as np

    from qvm_amd import circuits as C
    from qvm_amd import gates as G
    from qvm_amd.batch import qaoa_batch

    inst = C.qaoa_instances(args.instances, seed=4)
    circuits = [C.qaoa_maxcut(n, e, g, b) for n, e, g, b in inst]
    upd = sum(C.amp_updates(c, n) for c, (n, _, _, _) in zip(circuits, inst))
    out = {"metric": "QAOA batch: instances/s and amp-updates/s, one hipGraph launch per run",
           "instances": len(inst), "gates": sum(len(c) for c in circuits), "amp_updates_per_run": upd,
           "shots_per_instance": 1024, "results": {}}
    hcs = {}
    chosen = {}
    for dtype in ("f64", "f32"):
        t0 = time.perf_counter()
        # tile width picked by the autotuner (timed on every 4th instance)
        b, _, _ = qaoa_batch(inst, dtype, shots=1024, seed=9, use_graph=True, fma_tiles=True, tile_m="auto")
        t_plan = time.perf_counter() - t0
        chosen[dtype] = b.tile_m
        b.run()
        b.sync()
        t0 = time.perf_counter()
        reps = max(1, args.steps)
        for _ in range(reps):
            b.run()
        b.sync()
        dt = (time.perf_counter() - t0) / reps
        hc, shots = b.results()
        hcs[dtype] = hc
        out["results"][dtype] = {"s_per_run": dt, "instances_per_s": len(inst) / dt, "amp_updates_per_s": upd / dt,
                                 "plan_and_capture_s": t_plan, "launch_items": b.items,
                                 "tile_m": b.tile_m,
                                 "autotune_s_per_run_on_sample": {str(k): v for k, v in b.tile_timings.items()},
                                 "state_GiB": sum((1 << n) for n, _, _, _ in inst) * (16 if dtype == "f64" else 8) / 2 ** 30,
                                 "mean_expected_cut": float(np.mean([0.5 * (len(e) - h) for (n, e, _, _), h in zip(inst, hc)]))}
        log(f"batch {dtype}: {dt * 1e3:.1f} ms/run, {len(inst) / dt:.0f} instances/s, tile width {b.tile_m} "
            f"(autotuner sample: {', '.join(f'm={k} {v * 1e3:.1f} ms' for k, v in b.tile_timings.items())})")
        b.close()
    out["results"]["f32_max_abs_hc_diff_vs_f64"] = float(np.max(np.abs(hcs["f32"] - hcs["f64"])))
    # eager (uncaptured) FP64 run for the graph's launch-overhead saving
    b, _, _ = qaoa_batch(inst, "f64", shots=1024, seed=9, use_graph=False, fma_tiles=True, tile_m=chosen["f64"])
    b.run()
    b.sync()
    t0 = time.perf_counter()
    b.run()
    b.sync()
    out["results"]["f64_eager_s_per_run"] = time.perf_counter() - t0
    b.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.set_parallelization(19, args.cpu_threads)
        t0 = time.perf_counter()
        done, cu = 0, 0
        for c, (n, _, _, _) in zip(circuits, inst):
            psi = O.zero_state(n)
            for U, q in c:
                O.apply_compiled(psi, U, G.nat_tuple(*q))
            done += 1
            cu += C.amp_updates(c, n)
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cu / dt, "unit": "amp-updates/s", "cores": args.cpu_threads, "kind": "port",
                               "instances_per_s": done / dt,
                               "sample": f"first {done} instances in generation order, FP64 compiled mode, {dt:.1f} s"}
    print(json.dumps(out))
    return 0


def learned(args) -> int:
    """Config 5 in full: optionally retrain the cost model on launch timings
    this GPU records (--learn-samples launches per (n, dtype), n = 16..28,
    written to --save-model), then time EVERY candidate plan of the
    bench-scale mixed workload and report what the learned model, the
    analytic model and the fixed rule pick against the best candidate, with
    each picked state against the plain state."""
    from qvm_amd import circuits as C
    from qvm_amd import costmodel as CM

    out = {"metric": "learned plan choice vs fixed rule and analytic model, mixed bench-scale workload"}
    cm = CM.CostModel.shared()
    if args.save_model:
        t0 = time.perf_counter()
        recs = CM.collect([16, 20, 24, 28], ["f64", "f32"], per_config=args.learn_samples, seed=21)
        cm = CM.CostModel.trained(recs, epochs=4000)
        cm.save(args.save_model, recs)
        out["retrained"] = {"launches": len(recs), "s": time.perf_counter() - t0, "path": args.save_model}
    rep = CM.evaluate_planners(C.mixed_workload("bench"), cm, CM.AnalyticModel(), measure_all=True, log=log)
    tot = rep["totals_ms"]
    rep["speedup_vs_fixed_rule"] = tot["fixed_rule"] / tot["learned"]
    rep["speedup_vs_analytic"] = tot["analytic"] / tot["learned"]
    rep["share_of_best_gain"] = ((tot["fixed_rule"] - tot["learned"]) / (tot["fixed_rule"] - tot["best_candidate"])
                                 if tot["fixed_rule"] > tot["best_candidate"] else None)
    out.update(rep)
    print(json.dumps(out))
    return 0


def cpu_threads(requested: int):
    """Worker count for the oracle: the reference's worker pool has one
    worker per logical core (sysconf(_SC_NPROCESSORS_ONLN),
    src/utilities.lisp:202-239); a process confined by a cgroup CPU quota
    gets no more than its quota from extra workers, so the count is capped
    there.  Returns (threads, facts)."""
    from oracle import oracle as O
    logical = O.count_logical_cores()
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = logical
    threads = requested or min(x for x in (logical, quota, affinity) if x)
    return threads, {"logical_cores_sysconf": logical, "cgroup_quota_cores": quota, "affinity_cores": affinity}


def cpu_baseline(n, ga, gb, threads, cores_info):
    """The oracle (C restatement of the reference CPU path) at the step's
    full size n: one gate of each kind in the step (RX, RZ, CNOT from the
    layered circuit; H, CPHASE, SWAP from the QFT) is timed, and the step's
    gate mix is priced with those per-gate times.  Modes: compiled (the
    specialised matvec over complement groups, src/compile-gate.lisp:23-209,
    pdotimes) and interp-avx (the 1q/2q AVX kernels split by
    with-parallel-subdivisions, src/serial-kernels.lisp:176-267) on
    `threads` workers, and compiled on one thread (RX, RZ, CNOT only: the
    layered circuit's mix)."""
    import numpy as np
    from oracle import oracle as O
    from qvm_amd import gates as G
    def gate_kind(U):
        if not np.any(U - np.diag(np.diag(U))):
            return "RZ" if U.shape[0] == 2 else "CPHASE"
        if U.shape[0] == 2:
            return "H" if not np.any(U.imag) else "RX"
        if np.count_nonzero(U) == 4 and np.all(np.abs(U[U != 0]) == 1) and not np.any(U.imag):
            return "SWAP" if U[1, 2] == 1 else "CNOT"
        return "CPHASE"

    kinds = {}
    for U, q in ga + gb:
        kinds.setdefault(gate_kind(U), []).append((U, q))
    count = {k: len(v) for k, v in kinds.items()}
    upd = {k: G.changed_amplitudes(v[0][0], n) for k, v in kinds.items()}
    psi = O.zero_state(n)
    psi[:] = 2.0 ** (-n / 2)  # a dense state: every amplitude nonzero

    def per_gate(mode, names, workers):
        O.set_parallelization(19, workers)
        fn = {"compiled": O.apply_compiled, "interp-avx": O.apply_avx}[mode]
        t = {}
        for k in names:
            U, q = kinds[k][len(kinds[k]) // 2]
            t0 = time.perf_counter()
            fn(psi, U, G.nat_tuple(*q))
            t[k] = time.perf_counter() - t0
        return t

    modes = {}
    all_names = sorted(kinds)
    for mode, names, workers in (("compiled", all_names, threads), ("interp-avx", all_names, threads),
                                 ("compiled_1thread", ["RX", "RZ", "CNOT"], 1)):
        t = per_gate(mode.split("_")[0], [k for k in names if k in kinds], workers)
        priced = [k for k in t]
        step_s = sum(count[k] * t[k] for k in priced)
        step_upd = sum(count[k] * upd[k] for k in priced)
        modes[mode] = {"threads": workers, "per_gate_s": {k: round(v, 4) for k, v in t.items()},
                       "amp_updates_per_s": step_upd / step_s,
                       "priced_gates": {k: count[k] for k in priced}}
        log(f"cpu baseline {mode}: {step_upd / step_s:.3e} amp-updates/s ({workers} threads)")
    del psi
    O.set_parallelization(19, threads)
    best = max(("compiled", "interp-avx"), key=lambda m: modes[m]["amp_updates_per_s"])
    return {"value": modes[best]["amp_updates_per_s"], "unit": "amp-updates/s", "cores": threads, "kind": "port",
            "mode": best, "modes": modes, **cores_info,
            "sample": f"oracle at n={n}, FP64: one gate of each kind in the step "
                      f"({', '.join(f'{k} x{count[k]}' for k in all_names)}) timed once on a dense state; "
                      f"value = the step's amp-updates over its gate mix priced with those per-gate times "
                      f"(extrapolated, not the whole step run), faster of compiled / interp-avx on {threads} threads"}


def config1(threads, HipPureState):
    """The CPU-runnable configuration in full on both sides: 20 qubits, 20
    layered RX/RZ+CNOT layers (990 gates), then 10^4 measure-all shots of the
    prepared state (rule C(b) > p, the state reloaded per shot).  The GPU runs
    the plain path and the batched shot kernel; the oracle runs compiled mode
    on `threads` threads for the gates and the reference's bisection sampler
    for a bounded sample of the shots (scaled to 10^4)."""
    from oracle import oracle as O
    from qvm_amd import circuits as C
    from qvm_amd import gates as G
    import numpy as np
    n, shots = 20, 10_000
    gates = C.layered(n, 20, seed=1)
    nat = [(U, G.nat_tuple(*q)) for U, q in gates]
    u = C.uniforms(7, shots)
    st = HipPureState(n, "f64")
    st.sync()
    t0 = time.perf_counter()
    for U, q in nat:
        st.apply_matrix(U, q)
    st.sync()
    g_gates = time.perf_counter() - t0
    t0 = time.perf_counter()
    gshots = st.sample_measure_all(u)
    g_shots = time.perf_counter() - t0
    again, near = st.sample_ex(u, "gt", 1e-15)
    amps = st.amplitudes()
    st.close()
    O.set_parallelization(19, threads)
    psi = O.zero_state(n)
    t0 = time.perf_counter()
    for U, q in nat:
        O.apply_compiled(psi, U, q)
    c_gates = time.perf_counter() - t0
    k, t0 = 0, time.perf_counter()
    mism = 0
    while k < shots and time.perf_counter() - t0 < 3.0:
        b = O.sample_bisect(psi, float(u[k]))
        mism += int(b != int(gshots[k]))
        k += 1
    c_shot = (time.perf_counter() - t0) / max(k, 1)
    upd = C.amp_updates(gates, n)
    return {"n_qubits": n, "gates": len(gates), "shots": shots, "amp_updates": upd,
            "gpu_s_gates": g_gates, "gpu_s_shots": g_shots, "gpu_amp_updates_per_s": upd / g_gates,
            "cpu_s_gates": c_gates, "cpu_s_shots_scaled": c_shot * shots, "cpu_threads": threads,
            "cpu_shots_timed": k, "cpu_amp_updates_per_s": upd / c_gates,
            "state_max_abs_diff_gpu_vs_oracle": float(np.max(np.abs(amps - psi))),
            "shots_differing_in_timed_sample": mism,
            "near_boundary_1e-15": int(near.sum()),
            "flagged_rerun_equal": bool(np.array_equal(again, gshots))}


if __name__ == "__main__":
    sys.exit(main())
