"""tau_int / effective-sample estimators (nanomagmc_amd.stats) against closed
forms: white noise has tau_int = 1/2, an AR(1) process x_t = phi x_{t-1} +
e_t has rho(t) = phi^t and tau_int = (1 + phi) / (2 (1 - phi)).  CPU only."""
import math

import numpy as np
import pytest

from nanomagmc_amd import stats


def ar1(phi, n, rng):
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / math.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return x


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.9])
def test_tau_int_of_ar1(phi):
    rng = np.random.default_rng(int(phi * 10) + 1)
    x = ar1(phi, 200_000, rng)
    tau, W = stats.tau_int(x)
    want = (1 + phi) / (2 * (1 - phi))
    assert abs(tau - want) < 0.05 * want + 0.02, (tau, want)
    n_eff, tau2, closed = stats.ess(x)
    assert closed and tau2 == tau and abs(n_eff - x.size / (2 * tau)) < 1e-9


def test_autocorrelation_is_exact_on_a_known_series():
    x = np.array([1.0, -1.0, 1.0, -1.0])
    rho = stats.autocorr(x)
    assert np.allclose(rho, [1.0, -0.75, 0.5, -0.25])


def test_unclosed_window_is_flagged():
    x = np.cumsum(np.random.default_rng(0).normal(size=2000))   # random walk: no finite tau
    _, _, closed = stats.ess(x)
    assert not closed


def test_mean_err_matches_independent_case():
    rng = np.random.default_rng(5)
    x = rng.normal(size=100_000)
    m, e, tau = stats.mean_err(x)
    assert abs(m) < 5 * e and abs(e - 1 / math.sqrt(x.size)) < 0.05 / math.sqrt(x.size)


def test_batched_tau_equals_the_loop():
    from nanomagmc_amd.stats import ess, tau_int, tau_int_batch
    rng = np.random.default_rng(3)
    rows = []
    for phi in (0.0, 0.5, 0.9, 0.99):
        x = np.zeros(600)
        for t in range(1, 600):
            x[t] = phi * x[t - 1] + rng.normal()
        rows.append(x)
    rows.append(np.zeros(600))                       # constant series
    X = np.array(rows)
    tau, closed = tau_int_batch(X)
    for i, x in enumerate(rows):
        t, W = tau_int(x)
        assert abs(tau[i] - t) < 1e-12 * max(1.0, t), (i, tau[i], t)
        assert bool(closed[i]) == bool(ess(x)[2]), i


def test_multi_chain_tau_on_ar1_and_stuck_chains():
    # AR(1) chains x_n = phi x_{n-1} + e_n started stationary: tau_int =
    # (1 + phi) / (2 (1 - phi)) (1/2 for phi = 0); chains frozen at different
    # values count as about half a sample each; with one chain the estimate
    # agrees with the single-chain tau_int
    from nanomagmc_amd.stats import tau_int, tau_int_chains
    rng = np.random.default_rng(7)
    for phi in (0.0, 0.6, 0.9):
        M, N = 8, 20000
        x = np.zeros((M, N))
        x[:, 0] = rng.normal(size=M) / np.sqrt(1 - phi * phi)
        e = rng.normal(size=(M, N))
        for n in range(1, N):
            x[:, n] = phi * x[:, n - 1] + e[:, n]
        want = (1 + phi) / (2 * (1 - phi))
        tau, W, closed = tau_int_chains(x)
        assert closed and abs(tau - want) < 0.06 * want + 0.02, (phi, tau, want)
        t1, _ = tau_int(x[0])
        t1m, _, _ = tau_int_chains(x[:1])
        assert abs(t1m - t1) < 0.05 * t1 + 0.02, (phi, t1m, t1)
    stuck = np.repeat(np.array([[3.0], [5.0], [4.0], [7.0]]), 1000, axis=1)
    stuck[:, ::97] += 1.0          # rare one-sample flickers: the per-chain tau sees only these
    tau, W, closed = tau_int_chains(stuck)
    assert not closed and tau > 0.9 * 1000
    assert tau_int(stuck[0])[0] < 1.0
    # chains drifting apart (not mixing) are charged too
    drift = np.cumsum(rng.normal(size=(6, 4000)), axis=1)
    assert tau_int_chains(drift)[0] > 400
