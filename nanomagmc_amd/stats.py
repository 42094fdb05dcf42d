"""Autocorrelation statistics of sampled series: integrated autocorrelation
time with Sokal's automatic window, and tau_int-normalised effective sample
counts.  Plain NumPy on host arrays (series already copied off the device)."""
import math

import numpy as np


def autocorr(x):
    """Normalised autocorrelation rho(t), t = 0..n-1, via a zero-padded FFT."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    x = x - x.mean()
    v = np.dot(x, x) / n
    if n == 0 or v == 0:
        return np.ones(max(n, 1))
    f = np.fft.rfft(x, 2 * n)
    return np.fft.irfft(f * np.conj(f))[:n] / (n * v)


def tau_int(x, c=6.0):
    """tau = 1/2 + sum_{t=1}^{W} rho(t) with the smallest window W >= c tau(W)
    (Sokal).  Returns (tau, W); W == len(x) - 1 means no window closed and
    tau is a lower bound."""
    rho = autocorr(x)
    n = rho.size
    tau = 0.5
    for W in range(1, n):
        tau += rho[W]
        if W >= c * tau:
            return max(tau, 0.5), W
    return max(tau, 0.5), n - 1


def ess(x):
    """Effective sample count n / (2 tau_int) and whether the window closed
    well inside the series (W < n / 5)."""
    x = np.asarray(x)
    tau, W = tau_int(x)
    return x.size / (2.0 * tau), tau, W < x.size / 5


def mean_err(x):
    """Mean and its tau_int-corrected standard error."""
    x = np.asarray(x, dtype=np.float64)
    tau, _ = tau_int(x)
    return x.mean(), math.sqrt(2.0 * tau * x.var(ddof=1) / x.size), tau


def tau_int_batch(X, c=6.0):
    """tau_int of every row of X (chains x samples) at once, the same
    estimator as tau_int: tau(W) = 1/2 + sum_{t=1}^{W} rho(t) at the first
    W >= c tau(W).  Returns (tau, closed): closed is False where no window
    closed before the end of the series (tau is then a lower bound)."""
    X = np.asarray(X, dtype=np.float64)
    m, n = X.shape
    X = X - X.mean(axis=1, keepdims=True)
    v = (X * X).sum(axis=1) / n
    f = np.fft.rfft(X, 2 * n, axis=1)
    rho = np.fft.irfft(f * np.conj(f), axis=1)[:, :n]
    safe = np.where(v > 0, v, 1.0)
    rho = rho / (n * safe[:, None])
    rho[v == 0] = 1.0                                  # a constant series counts as frozen, as in autocorr
    cum = 0.5 + np.cumsum(rho[:, 1:], axis=1)          # tau(W), W = 1 .. n-1
    W = np.arange(1, n)[None, :]
    ok = W >= c * cum
    has = ok.any(axis=1)
    first = ok.argmax(axis=1)
    tau = np.where(has, cum[np.arange(m), first], cum[:, -1])
    tau = np.maximum(tau, 0.5)
    return tau, has & (np.where(has, first + 1, n) < n / 5)


def tau_int_chains(X, c=6.0):
    """tau_int of M chains of one observable together (X: chains x samples),
    with the multi-chain autocorrelation of Gelman et al., Bayesian Data
    Analysis 3rd ed., sec. 11.5:
        W = mean within-chain variance, B = N var(chain means),
        var+ = (N - 1) / N W + B / N,
        V_t = mean over chains and n of (x_n - x_{n-t})^2   (the variogram),
        rho(t) = 1 - V_t / (2 var+),
    and Sokal's window on tau(W) = 1/2 + sum_{t=1}^{W} rho(t), as in tau_int.
    Chains stuck at different values have W ~ 0 and B > 0, so rho stays near
    1 and the M chains count as about M / 2 samples: the estimator charges
    chains that do not mix, which the per-chain tau_int cannot see.
    Returns (tau, W, closed); the effective sample count is M N / (2 tau)."""
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[None, :]
    M, N = X.shape
    Xc = X - X.mean(axis=1, keepdims=True)
    Wv = (Xc * Xc).sum(axis=1) / max(N - 1, 1)
    Wm = Wv.mean()
    B = N * X.mean(axis=1).var(ddof=1) if M > 1 else 0.0
    var_plus = (N - 1) / N * Wm + B / N
    if var_plus <= 0:
        return float(N - 0.5), N - 1, False      # every chain constant at one value: nothing to estimate
    # variogram: sum_n (x_n - x_{n-t})^2 = S2_hi(t) + S2_lo(t) - 2 C(t)
    f = np.fft.rfft(Xc, 2 * N, axis=1)
    C = np.fft.irfft(f * np.conj(f), axis=1)[:, :N].sum(axis=0)
    sq = (Xc * Xc)
    cum = np.concatenate([[0.0], np.cumsum(sq.sum(axis=0))])
    t = np.arange(1, N)
    hi = cum[N] - cum[t]           # sum_{n >= t} x_n^2
    lo = cum[N - t]                # sum_{n < N - t} x_n^2
    V = (hi + lo - 2.0 * C[1:]) / (M * (N - t))
    rho = 1.0 - V / (2.0 * var_plus)
    tau = 0.5
    for k, r in enumerate(rho, start=1):
        tau += r
        if k >= c * tau:
            return max(float(tau), 0.5), k, k < N / 5
    return max(float(tau), 0.5), N - 1, False
