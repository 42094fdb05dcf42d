"""Closed forms and exact enumerations the parity tests check against.

- Exact enumeration of small periodic Ising lattices (all 2^N states).
- Kaufman's exact partition function of the L x L periodic Ising model
  (B. Kaufman, Phys. Rev. 76, 1232 (1949); the form of Ferdinand & Fisher,
  Phys. Rev. 185, 832 (1969)), pinned against enumeration at L = 4.
- Onsager's infinite-lattice energy per site.
- Langevin function for classical Heisenberg dimers / single spins in a field.
- Sokal's automatic-windowing integrated autocorrelation time.
"""
import math

import numpy as np


def ising_periodic_energies(Lx, Ly):
    """E(state) = -sum_<ij> s_i s_j over all 2^(Lx*Ly) states of a periodic
    Lx x Ly lattice (J = 1, h = 0), with M(state)."""
    N = Lx * Ly
    states = np.arange(2 ** N, dtype=np.int64)
    bits = ((states[:, None] >> np.arange(N)[None, :]) & 1).astype(np.int8) * 2 - 1
    s = bits.reshape(-1, Lx, Ly)
    E = -(s * np.roll(s, 1, axis=1)).sum(axis=(1, 2)) - (s * np.roll(s, 1, axis=2)).sum(axis=(1, 2))
    M = s.sum(axis=(1, 2))
    return E.astype(np.float64), M.astype(np.float64)


def enum_averages(E, M, beta):
    w = -beta * E
    w = np.exp(w - w.max())
    Z = w.sum()
    e = (w * E).sum() / Z
    e2 = (w * E * E).sum() / Z
    am = (w * np.abs(M)).sum() / Z
    m2 = (w * M * M).sum() / Z
    return {"E": e, "C": beta * beta * (e2 - e * e), "absM": am, "M2": m2}


def _logaddexp_signed(terms):
    # terms: list of (sign, log|x|)
    mx = max(t[1] for t in terms)
    s = sum(sg * math.exp(lg - mx) for sg, lg in terms)
    return mx + math.log(s)


def kaufman_logZ(L, K):
    """ln Z of the L x L periodic Ising model at coupling K = beta J."""
    c2, s2 = math.cosh(2 * K), math.sinh(2 * K)
    ch = c2 * c2 / s2   # cosh 2K coth 2K

    def gamma(l):
        if l == 0:
            return 2 * K + math.log(math.tanh(K))   # may be negative (T < Tc)
        return math.acosh(ch - math.cos(math.pi * l / L))

    odd = [gamma(2 * r + 1) for r in range(L)]
    even = [gamma(2 * r) for r in range(L)]

    def logprod_cosh(gs):
        return sum(math.log(2 * math.cosh(L * g / 2)) for g in gs)

    def logprod_sinh(gs):
        sgn, tot = 1.0, 0.0
        for g in gs:
            v = 2 * math.sinh(L * g / 2)
            if v == 0:
                return 0.0, -math.inf
            sgn *= 1.0 if v > 0 else -1.0
            tot += math.log(abs(v))
        return sgn, tot

    terms = [(1.0, logprod_cosh(odd))]
    sg, lg = logprod_sinh(odd)
    if lg > -math.inf:
        terms.append((sg, lg))
    terms.append((1.0, logprod_cosh(even)))
    sg, lg = logprod_sinh(even)
    if lg > -math.inf:
        terms.append((sg, lg))
    return math.log(0.5) + 0.5 * L * L * math.log(2 * s2) + _logaddexp_signed(terms)


def kaufman_energy(L, T, h=1e-5):
    """<E> (total, J = 1) of the L x L periodic Ising model at temperature T."""
    b = 1.0 / T
    return -(kaufman_logZ(L, b + h) - kaufman_logZ(L, b - h)) / (2 * h)


def onsager_energy_per_site(T):
    """Onsager's u(T) for the infinite square lattice (J = 1)."""
    from scipy import integrate
    K = 1.0 / T
    k = 2 * math.sinh(2 * K) / math.cosh(2 * K) ** 2
    kp = 2 * math.tanh(2 * K) ** 2 - 1
    K1, _ = integrate.quad(lambda t: 1 / math.sqrt(1 - k * k * math.sin(t) ** 2), 0, math.pi / 2)
    return -(1 / math.tanh(2 * K)) * (1 + (2 / math.pi) * kp * K1)


def langevin(x):
    return 1 / math.tanh(x) - 1 / x


def dmi_dimer_energy(beta, J, D):
    """<E> of two classical unit spins with E = -J s1.s2 - D.(s1 x s2),
    |D| = D.  Integrating s1 over the sphere against the field
    b = J s2 + s2 x D, |b|^2 = J^2 + D^2 (1 - u^2) with u = cos(s2, D), leaves
    Z ~ int_{-1}^{1} sinh(beta K) / (beta K) du, K = |b|; <E> = -d ln Z / d beta."""
    from scipy import integrate

    def K(u):
        return math.sqrt(J * J + D * D * (1 - u * u))

    def f(u):
        x = beta * K(u)
        return math.sinh(x) / x

    def df(u):
        x = beta * K(u)
        return (math.cosh(x) - math.sinh(x) / x) / beta

    Z, _ = integrate.quad(f, -1, 1, epsabs=0, epsrel=1e-12)
    dZ, _ = integrate.quad(df, -1, 1, epsabs=0, epsrel=1e-12)
    return -dZ / Z


def tau_int(x, c=6.0):
    """Sokal's automatic windowing: tau = 1/2 + sum_{t=1}^{W} rho(t), with the
    smallest W >= c * tau(W)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    x = x - x.mean()
    v = np.dot(x, x) / n
    if v == 0:
        return 0.5
    f = np.fft.rfft(np.concatenate([x, np.zeros(n)]))
    acf = np.fft.irfft(f * np.conj(f))[:n] / (n * v)
    tau = 0.5
    for W in range(1, n):
        tau += acf[W]
        if W >= c * tau:
            break
    return max(tau, 0.5)


def mean_err(x):
    """Mean and its tau_int-corrected standard error."""
    x = np.asarray(x, dtype=np.float64)
    t = tau_int(x)
    return x.mean(), math.sqrt(2 * t * x.var(ddof=1) / x.size), t
