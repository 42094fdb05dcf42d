"""Lattice fixtures of the reference's gtest suite, rebuilt with the oracle.

gen_fm / gen_afm / gen_skyrm restate Old/tests/test_functions.hpp:19-128
(edge 10, interaction file 3d_nn_bloch.dat, intended row-major site order).
"""
import numpy as np


def box_locs(dim, edge):
    """Row-major box scan, last coordinate fastest (test_functions.hpp:25-37)."""
    g = np.indices((edge,) * dim).reshape(dim, -1).T
    out = np.zeros((g.shape[0], 3), dtype=np.int32)
    out[:, :dim] = g
    return out


def nn2d():
    """2-D nearest-neighbour table (+y, -y, +x, -x), J = 1, D = 0."""
    d = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0]], dtype=np.int32)
    return d, np.ones(4), np.zeros((4, 3))


def neel2d(D=1.0):
    """2-D interfacial (Neel) DMI: D_ij = D z x r_ij for the four bonds."""
    d, J, _ = nn2d()
    Dv = np.zeros((4, 3))
    for k, (dx, dy, _) in enumerate(d):
        r = np.array([dx, dy, 0.0])
        Dv[k] = D * np.cross([0.0, 0.0, 1.0], r)
    return d, J, Dv


def up_spins(n, ising):
    s = np.zeros((n, 3))
    s[:, 0 if ising else 2] = 1.0
    return s


def afm_spins(locs, dim, ising):
    s = up_spins(len(locs), ising)
    odd = (locs[:, :dim].sum(axis=1) % 2) == 1
    s[odd] *= -1.0
    return s


def skyrm_spins(edge=30, dim=3):
    """gen_skyrm (test_functions.hpp:79-128): a skyrmion of radius mid-3 in
    every z-slice; spins outside the radius are the zero vector.  float32
    arithmetic as in the fixture."""
    locs = box_locs(dim, edge)
    mid = edge // 2
    sk_R = np.float32(mid - 3)
    k = np.float32(np.pi / float(sk_R))
    s = np.zeros((len(locs), 3))
    for i, p in enumerate(locs):
        dx = np.float32(p[0] - mid)
        dy = np.float32(p[1] - mid)
        r = np.float32(np.sqrt(np.float64(dx) ** 2 + np.float64(dy) ** 2))
        if r >= sk_R:
            continue
        phi = np.float32(np.arctan2(dx, dy))
        kr = np.float32(k * r)
        s[i, 0] = np.float32(-np.float32(np.sin(kr)) * np.float32(np.cos(phi)))
        s[i, 1] = np.float32(np.float32(np.sin(kr)) * np.float32(np.sin(phi)))
        s[i, 2] = np.float32(np.cos(kr))
    return locs, s


def random_unit(n, rng):
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)
