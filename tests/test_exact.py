"""Pins the closed forms in tests/exact.py that the GPU statistics tests
compare against (CPU only)."""
import math

import numpy as np

from tests import exact


def test_dmi_dimer_reduces_to_langevin_without_dmi():
    # D = 0: <E> = -J L(beta J) for an isotropic pair
    for beta in (0.3, 1.0, 2.5):
        assert abs(exact.dmi_dimer_energy(beta, 1.0, 0.0) + exact.langevin(beta)) < 1e-10


def test_dmi_dimer_energy_is_minus_dlnz_dbeta():
    # the analytic derivative inside dmi_dimer_energy against a central
    # difference of ln Z over beta
    from scipy import integrate

    def lnZ(beta, J=1.0, D=0.8):
        f = lambda u: math.sinh(beta * math.sqrt(J * J + D * D * (1 - u * u))) / (
            beta * math.sqrt(J * J + D * D * (1 - u * u)))
        return math.log(integrate.quad(f, -1, 1, epsabs=0, epsrel=1e-13)[0])

    for beta in (0.5, 2.0):
        h = 1e-5
        fd = -(lnZ(beta + h) - lnZ(beta - h)) / (2 * h)
        assert abs(fd - exact.dmi_dimer_energy(beta, 1.0, 0.8)) < 1e-6


def test_dmi_dimer_energy_matches_direct_sphere_average():
    # brute force: uniform pairs of unit vectors weighted by exp(-beta E)
    rng = np.random.default_rng(3)
    n = 2_000_000
    s = rng.normal(size=(2, n, 3))
    s /= np.linalg.norm(s, axis=2, keepdims=True)
    D = np.array([0.3, -0.4, 0.5])
    E = -np.einsum("ij,ij->i", s[0], s[1]) - np.cross(s[0], s[1]) @ D
    for beta in (0.5, 1.5):
        w = np.exp(-beta * (E - E.min()))
        want = float((w * E).sum() / w.sum())
        assert abs(want - exact.dmi_dimer_energy(beta, 1.0, float(np.linalg.norm(D)))) < 5e-3
        # and the DMI matters at this size: dropping it moves <E> well beyond that
        assert abs(want - exact.dmi_dimer_energy(beta, 1.0, 0.0)) > 3e-2
