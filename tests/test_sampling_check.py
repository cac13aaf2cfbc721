"""CPU checks of the sampling-parity checker itself: it passes the oracle
against itself, counts draws placed on CDF values, and rejects a planted
off-by-one away from a boundary and a near flag on a draw far from every
boundary."""
import numpy as np
import pytest

from tests.sampling_check import check_draws, pair_bound


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def test_checker_accepts_oracle_and_counts_boundaries(oracle):
    rng = np.random.default_rng(3)
    psi = rand_state(12, rng)
    cdf = oracle.cdf(psi)
    u = np.concatenate([rng.random(500), cdf[:50]])
    want = oracle.sample_multiple(psi, u)
    r = check_draws(cdf, u, want, want, 1e-15)
    assert r["mismatch"] == 0 and r["near_oracle"] >= 50
    # the other tie rule differs only on the draws placed on CDF values,
    # and those are within the bound (distance 0)
    gt = np.array([oracle.cdf_search(cdf, np.nextafter(p, 2.0)) for p in u], dtype=np.uint64)
    r = check_draws(cdf, u, gt, want, 1e-15)
    assert 0 < r["mismatch"] <= 50 and r["mismatch_beyond_eps"] == 0


def test_checker_rejects_planted_errors(oracle):
    rng = np.random.default_rng(4)
    psi = rand_state(10, rng)
    cdf = oracle.cdf(psi)
    u = rng.random(200)
    want = oracle.sample_multiple(psi, u)
    # an index one too high for a draw in the middle of its interval
    j = int(np.argmax(np.minimum(np.abs(u - cdf[want]), np.abs(u - np.where(want > 0, cdf[want - 1], 0)))))
    bad = want.copy()
    bad[j] = min(int(bad[j]) + 1, cdf.size - 1) if int(bad[j]) + 1 < cdf.size else int(bad[j]) - 1
    with pytest.raises(AssertionError):
        check_draws(cdf, u, bad, want)
    flags = np.zeros(u.size, dtype=bool)
    flags[j] = True
    with pytest.raises(AssertionError):
        check_draws(cdf, u, want, want, 1e-15, flags)


def test_pair_bound_is_the_sequential_error_scale():
    cdf = np.linspace(1.0 / 1024, 1.0, 1024)
    b = pair_bound(cdf, np.array([0, 1, 1023]))
    assert b[0] == 0.0
    assert abs(b[2] - 2 * 1023 * 2.0 ** -53) < 1e-20
