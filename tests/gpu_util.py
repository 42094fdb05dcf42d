"""Helpers shared by the GPU parity tests (construct matching GPU contexts
and oracle fields from one description)."""
import math

import numpy as np

from tests.fixtures import box_locs


def engine_bonds(d, J, D):
    from nanomagmc_amd import Bond
    return [Bond(d[k], J[k], D[k]) for k in range(len(J))]


def mask_from_locs(locs, dim, n_box, origin=(0, 0, 0)):
    m = np.zeros(n_box[:dim], dtype=np.uint8)
    rel = locs[:, :dim] - np.asarray(origin[:dim])[None, :]
    m[tuple(rel.T)] = 1
    return m.reshape(-1)


def make_pair(orc, dim, edge, periodic, bd, J, D, ising, use_d=False, locs=None, trial="sphere",
              seed=0x4E4D4301, ninst=1, n_box=None, origin=(0, 0, 0)):
    """A GPU context with `ninst` copies of one lattice and the oracle Field."""
    from nanomagmc_amd import Lattices
    full = locs is None
    if full:
        locs = box_locs(dim, edge)
    if n_box is None:
        n_box = (edge,) * dim
    inst = {"n": tuple(n_box), "origin": tuple(origin), "edge": edge}
    if not full:
        inst["mask"] = mask_from_locs(locs, dim, n_box, origin)
    lat = Lattices(dim, ising, periodic, engine_bonds(bd, J, D), [dict(inst) for _ in range(ninst)],
                   use_j=True, use_d=use_d, trial=trial, seed=seed)
    f = orc.Field(locs, dim, edge, periodic, bd, J, D)
    return lat, f


def replica_stats(per_replica):
    """Mean over replicas and its standard error."""
    x = np.asarray(per_replica, dtype=np.float64)
    return x.mean(), x.std(ddof=1) / math.sqrt(x.size)


def estimators(E, M, beta, N):
    """Per-replica <E>, <|M|>, C = beta^2 var(E) / N, chi = beta (<M^2> - <|M|>^2) / N."""
    aM = np.abs(M) if M.ndim == 1 else np.linalg.norm(M, axis=-1)
    return {"E": E.mean(), "absM": aM.mean(), "C": beta * beta * E.var() / N,
            "chi": beta * ((aM * aM).mean() - aM.mean() ** 2) / N}


def oracle_chain(orc, f, H, beta, nsamples, sweeps, rng, burn=200):
    """The reference's random-site Metropolis on the host in FP64: uniform
    sphere trials, accept iff dE <= 0 or -dE*beta > log(u).  Returns the
    E_s, M and Q_wrap series."""
    n = f.n
    s = rng.normal(size=(n, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)

    def attempts(k):
        z = rng.uniform(-1, 1, k)
        ph = rng.uniform(0, 2 * math.pi, k)
        r = np.sqrt(1 - z * z)
        t = np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)
        return rng.integers(0, n, k), t, rng.uniform(size=k)

    for _ in range(burn):
        f.equil_replay(s, H, beta, *attempts(sweeps * n), use_d=True)
    E = np.zeros(nsamples)
    M = np.zeros((nsamples, 3))
    Q = np.zeros(nsamples)
    for k in range(nsamples):
        f.equil_replay(s, H, beta, *attempts(sweeps * n), use_d=True)
        E[k] = f.energy(s, H, use_d=True, convention=1)
        M[k] = f.mag(s)
        Q[k] = f.tcharge(s, wrap=True)[0]
    return E, M, Q
