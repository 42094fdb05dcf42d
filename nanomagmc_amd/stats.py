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
