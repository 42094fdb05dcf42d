"""Pins the network oracle (oracle/netref.py): its analytic gradient against
central finite differences in FP64, and the vMF log-density against
numerical integration over the sphere.  CPU only."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def netref():
    from oracle import netref
    return netref


def random_case(netref, rng, n=2):
    p = netref.init_params(rng)
    patches = np.zeros((n, 24, 24, 4))
    s = rng.normal(size=(n, 24, 24, 3))
    s /= np.linalg.norm(s, axis=-1, keepdims=True)
    patches[..., :3] = s
    patches[..., 3] = 1.0
    patches[:, 4:20, 4:20, :] = 0.0
    t = rng.normal(size=(n, 256, 3))
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    return p, patches, t


def test_gradient_matches_finite_differences(netref):
    rng = np.random.default_rng(0)
    p, patches, t = random_case(netref, rng)
    loss, g = netref.loss_grad(p, patches, t)
    assert g.size == netref.n_params() == p.size
    # probe coordinates in every layer's weights and biases
    offs, idx = 0, []
    for l, c in enumerate(netref.CIN):
        K = 9 * c
        for _ in range(6):
            k = int(rng.integers(K))
            if l == 0:
                k = (k // c) * c + int(rng.integers(4))
            n = int(rng.integers(netref.NOUT[l]))
            idx.append(offs + k * 32 + n)
        idx.append(offs + K * 32 + int(rng.integers(netref.NOUT[l])))
        offs += K * 32 + 32
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (netref.loss_grad(pp, patches, t)[0] - netref.loss_grad(pm, patches, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_vmf_density_normalised(netref):
    # integral over S^2 of q = 1, and the mean cosine = coth k - 1/k
    for kappa in (0.05, 0.7, 5.0, 40.0):
        th = np.linspace(0, np.pi, 200001)
        mu = np.array([0.0, 0.0, 1.0])
        x = np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], -1)
        lq = netref.logq(mu[None, :], np.full(th.shape, kappa), x)
        f = np.exp(lq) * 2 * np.pi * np.sin(th)
        assert abs(np.trapezoid(f, th) - 1) < 1e-6
        mc = np.trapezoid(f * np.cos(th), th)
        assert abs(mc - (1 / math.tanh(kappa) - 1 / kappa)) < 1e-6
