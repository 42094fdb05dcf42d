"""Golden vectors of the oracle, committed as tests/golden/vectors.npz.

Written by the oracle (oracle/oracle.py, the FP64 restatement of the
reference's thermodynamics.cpp and state.cpp), not by the reference itself.
Each array set is one case:

* de_<case>_*: dE_func (thermodynamics.cpp:10-69) for 64 (site, trial)
  pairs on one seeded configuration: Ising 2-D periodic, Ising 3-D open,
  Heisenberg 3-D Bloch periodic, Heisenberg 2-D Neel open; with the
  interaction table (dirs, J, D), the field h along the spin axis of the
  model, the spins, and for Heisenberg the site's scale |H + V| |s - t|.
* rp_ising2d_*: the reference's single-site update (state.cpp:190-201)
  fed 4000 (site, trial, uniform k / 2^32) steps on a 16 x 16 periodic Ising
  lattice; the accept decisions.
* tc_skyrm30: the topological charge per z-slice (thermodynamics.cpp:165-221)
  of the 30^3 skyrmion fixture of the reference's gtests.

Run `python tests/golden/make_golden.py` to rewrite the file;
tests/test_golden.py regenerates every array and checks it against the file.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from tests.fixtures import box_locs, neel2d, nn2d, random_unit, skyrm_spins  # noqa: E402

OUT = os.path.join(HERE, "vectors.npz")
BLOCH = os.path.join(HERE, "Js", "3d_nn_bloch.dat")

# name: dim, edge, periodic, ising, use_d, table
DE_CASES = {
    "ising2d": (2, 12, 1, True, False, "nn2d"),
    "ising3d": (3, 6, 0, True, False, "bloch0"),
    "bloch3d": (3, 6, 1, False, True, "bloch"),
    "neel2d": (2, 12, 0, False, True, "neel"),
}


def table(orc, name):
    if name == "nn2d":
        return nn2d()
    if name == "neel":
        return neel2d(0.3)
    return orc.parse_bonds(BLOCH, 3, 1.0, 0.0 if name == "bloch0" else 0.3)


def de_case(orc, name):
    dim, edge, periodic, ising, use_d, tab = DE_CASES[name]
    rng = np.random.default_rng(0x4E4D47 + sum(map(ord, name)))
    bd, J, D = table(orc, tab)
    locs = box_locs(dim, edge)
    f = orc.Field(locs, dim, edge, periodic, bd, J, D)
    h = 0.1
    H = [h, 0.0, 0.0] if ising else [0.0, 0.0, h]
    n = 64
    if ising:
        s = np.zeros((f.n, 3))
        s[:, 0] = rng.choice([-1.0, 1.0], f.n)
        trials = np.zeros((n, 3))
        trials[:, 0] = rng.choice([-1.0, 1.0], n)
    else:   # fp32 values, as the device holds them
        s = random_unit(f.n, rng).astype(np.float32).astype(np.float64)
        trials = random_unit(n, rng).astype(np.float32).astype(np.float64)
    sites = rng.integers(0, f.n, n).astype(np.int32)
    want = np.array([f.delta_e(s, int(sites[i]), trials[i], H, use_d=use_d) for i in range(n)])
    out = {"dirs": np.asarray(bd, dtype=np.int32), "J": np.asarray(J, dtype=np.float64),
           "D": np.asarray(D, dtype=np.float64), "h": np.float64(h), "spins": s, "sites": sites,
           "trials": trials, "want": want}
    if not ising:
        scale = np.zeros(n)
        for i in range(n):
            j = int(sites[i])
            V = np.array(H, dtype=np.float64)
            for c in range(f.ncount[j]):
                k, nb = f.nchoice[j, c], s[f.neigh[j, c]]
                V += f.J[k] * nb + (np.cross(nb, f.D[k]) if use_d else 0.0)
            scale[i] = np.linalg.norm(V) * np.linalg.norm(s[j] - trials[i])
        out["scale"] = scale
    return out


def replay_case(orc):
    rng = np.random.default_rng(0x4E4D4752)
    bd, J, D = nn2d()
    edge = 16
    f = orc.Field(box_locs(2, edge), 2, edge, 1, bd, J, D)
    s = np.zeros((f.n, 3))
    s[:, 0] = rng.choice([-1.0, 1.0], f.n)
    T, h, n = 2.3, 0.1, 4000
    sites = rng.integers(0, f.n, n).astype(np.int32)
    tr = rng.choice([-1, 1], n).astype(np.int8)
    k = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    trials = np.zeros((n, 3))
    trials[:, 0] = tr
    acc = f.equil_replay(s.copy(), [h, 0.0, 0.0], 1.0 / T, sites, trials, k / 4294967296.0)   # mutates its copy
    return {"spins": s, "T": np.float64(T), "h": np.float64(h), "sites": sites, "trials": tr, "k": k,
            "accept": np.asarray(acc, dtype=np.uint8)}


def charge_case(orc):
    locs, s = skyrm_spins(30, 3)
    bd, J, D = orc.parse_bonds(BLOCH, 3, 1.0, 0.3)
    f = orc.Field(locs, 3, 30, 0, bd, J, D)
    return np.asarray(f.tcharge(s), dtype=np.float64)


def vectors(orc):
    """Every golden array, by name."""
    out = {}
    for name in DE_CASES:
        for key, v in de_case(orc, name).items():
            out["de_%s_%s" % (name, key)] = v
    for key, v in replay_case(orc).items():
        out["rp_ising2d_%s" % key] = v
    out["tc_skyrm30"] = charge_case(orc)
    return out


if __name__ == "__main__":
    from oracle import oracle
    oracle.build()
    v = vectors(oracle)
    np.savez_compressed(OUT, **v)
    print("wrote %s (%d arrays)" % (OUT, len(v)))
