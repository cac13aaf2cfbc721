"""Sampling parity with an explicit FP64 bound.

The engine and the oracle form the cumulative probabilities C(i) =
sum_{k<=i} |psi_k|^2 in different orders: the oracle sequentially
(cumulative-distribution-function, src/wavefunction.lisp:380-395) or as the
bisection's range sums (src/measurement.lisp:179-225), the engine as a
chunked tree.  Every order forms C(i) from its i+1 nonnegative terms with i
additions, and the sum of nonnegative terms so formed is within
gamma_i * S(i) of the exact S(i), gamma_i = i*u / (1 - i*u), u = 2^-53.  Two
such values differ by at most 2 * gamma_i * S(i).

A draw p that two samplers resolve differently therefore sits at a
boundary: with i the smaller of the two indices, |p - C(i)| <= 2 * gamma_i *
S(i) (for the rule C >= p: the engine's C(i) >= p > the oracle's C(i), or the
reverse; the same for C > p).  `check_draws` asserts exactly that for every
mismatch, and counts the draws within eps of a boundary.
"""
from __future__ import annotations

import numpy as np

UNIT = 2.0 ** -53


def gamma(i):
    i = np.asarray(i, dtype=np.float64)
    return i * UNIT / (1.0 - i * UNIT)


def pair_bound(cdf: np.ndarray, i) -> np.ndarray:
    """Largest difference two summation orders can give for C(i)."""
    i = np.asarray(i, dtype=np.int64)
    g = gamma(i)
    s = cdf[i] * (1.0 + g)  # bound on the exact S(i) from the oracle's value
    return 2.0 * g * s


def near_mask(cdf: np.ndarray, u: np.ndarray, idx: np.ndarray, eps: float) -> np.ndarray:
    """Draws within eps of C(idx) or C(idx - 1) under the oracle's CDF."""
    idx = np.asarray(idx, dtype=np.int64)
    hi = cdf[idx]
    lo = np.where(idx > 0, cdf[np.maximum(idx - 1, 0)], 0.0)
    return (np.abs(u - hi) < eps) | (np.abs(u - lo) < eps)


def check_draws(cdf: np.ndarray, u, got, want, eps: float = 1e-15, engine_near=None) -> dict:
    """Asserts every mismatch between engine (got) and oracle (want) lies
    within the pair bound of the oracle's C at the smaller index; when the
    engine's near flags are given, asserts each flagged draw is within eps
    plus that bound of C(b) or C(b-1) of the engine's index b.  Returns the
    counts."""
    u = np.asarray(u, dtype=np.float64)
    got = np.asarray(got, dtype=np.int64)
    want = np.asarray(want, dtype=np.int64)
    bad = np.nonzero(got != want)[0]
    for j in bad:
        i = int(min(got[j], want[j]))
        b = float(pair_bound(cdf, i))
        assert abs(u[j] - cdf[i]) <= b, (
            f"draw {j}: u={u[j]!r} engine {got[j]} oracle {want[j]}: |u - C({i})| = "
            f"{abs(u[j] - cdf[i]):.3e} exceeds the summation bound {b:.3e}")
    near_or = near_mask(cdf, u, want, eps)
    out = {
        "draws": int(u.size),
        "mismatch": int(bad.size),
        "mismatch_beyond_eps": int(np.count_nonzero(~near_mask(cdf, u[bad], want[bad], eps))) if bad.size else 0,
        "near_oracle": int(np.count_nonzero(near_or)),
    }
    if engine_near is not None:
        engine_near = np.asarray(engine_near, dtype=bool)
        fl = np.nonzero(engine_near)[0]
        if fl.size:
            slack = eps + pair_bound(cdf, got[fl]) + pair_bound(cdf, np.maximum(got[fl] - 1, 0))
            hi = cdf[got[fl]]
            lo = np.where(got[fl] > 0, cdf[np.maximum(got[fl] - 1, 0)], 0.0)
            ok = (np.abs(u[fl] - hi) <= slack) | (np.abs(u[fl] - lo) <= slack)
            assert ok.all(), f"engine flagged draws far from every boundary: {fl[~ok][:5]}"
        out["near_engine"] = int(fl.size)
    return out
