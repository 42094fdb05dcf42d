"""Sampled observables of the device sampler against exact answers and the
FP64 oracle, each with its stated error (3 sigma, errors from the spread of
independent replicas, so tau_int is accounted for):

* 2-D Ising 4x4 periodic: <E>, <|M|>, C and chi against exact enumeration
  of all 2^16 states;
* 2-D Heisenberg 6x6 periodic with Neel DMI and a field: <E_s>, <|M|>, C,
  chi and <|Q|> (plaquette charge with periodic wrap) against the oracle's
  random-site Metropolis chain (the reference's state::equil with
  dE_func, Old/lib/state.cpp:171-203), run in FP64 on the host;
* 3-D Heisenberg 4^3 periodic with Bloch DMI and a field, sampled with the
  benchmark's cone trials (half-angles adapted on the device, then frozen):
  <E_s>, <|M|>, C and chi against the same oracle chain (uniform-sphere
  trials).

Everything is seeded: device sweeps draw from Philox counters and the
reductions have a fixed order, so the outcome is the same on every run.
Needs an MI355X."""
import math

import numpy as np
import pytest

from tests import exact
from tests.fixtures import box_locs, neel2d
from tests.gpu_util import estimators, oracle_chain, replica_stats

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("T", [1.8, 2.4, 3.5])
def test_ising_4x4_observables_match_exact_enumeration(T):
    from nanomagmc_amd import Bond, Lattices
    L, R, ns = 4, 24, 4000
    bonds = [Bond((0, 1), 1.0), Bond((0, -1), 1.0), Bond((1, 0), 1.0), Bond((-1, 0), 1.0)]
    lat = Lattices(2, True, True, bonds, [{"n": (L, L), "edge": L}] * R, seed=0x4E4D4311 + int(10 * T))
    beta = 1.0 / T
    lat.set_params([beta] * R, 0.0)
    lat.randomize()
    lat.equil(500)
    lat.series_reset(ns, with_charge=False)
    lat.run_samples(ns, 3)
    obs, _, _ = lat.get_series()
    E_all, M_all = exact.ising_periodic_energies(L, L)
    ex = exact.enum_averages(E_all, M_all, beta)
    want = {"E": ex["E"], "absM": ex["absM"], "C": ex["C"] / (L * L),
            "chi": beta * (ex["M2"] - ex["absM"] ** 2) / (L * L)}
    per = [estimators(obs[i, :, 9], obs[i, :, 0], beta, L * L) for i in range(R)]
    for k in want:
        m, err = replica_stats([p[k] for p in per])
        print("T=%.2f %-5s device %.5f +- %.5f exact %.5f" % (T, k, m, err, want[k]))
        assert abs(m - want[k]) < 3 * err, (T, k, m, err, want[k])


def test_heisenberg_dmi_observables_match_oracle_chain(orc):
    from nanomagmc_amd import Bond, Lattices
    L, D, h, T = 6, 0.8, 0.3, 1.0
    beta = 1.0 / T
    d, Jv, Dv = neel2d(D)
    R, ns = 16, 8000
    lat = Lattices(2, False, True, [Bond(d[k], Jv[k], Dv[k]) for k in range(4)], [{"n": (L, L), "edge": L}] * R,
                   use_d=True, seed=0x4E4D4312)
    lat.set_params([beta] * R, h)
    lat.randomize()
    lat.equil(500)
    lat.series_reset(ns, with_charge=True)
    lat.run_samples(ns, 5)
    obs, _, qw = lat.get_series()
    N = L * L
    dev = []
    for i in range(R):
        st = estimators(obs[i, :, 10], obs[i, :, 0:3], beta, N)
        st["absQ"] = np.abs(qw[i, :, 0]).mean()
        dev.append(st)
    f = orc.Field(box_locs(2, L), 2, L, 1, d, Jv, Dv)
    rng = np.random.default_rng(0x4E4D4313)
    ref = []
    for _ in range(8):
        E, M, Q = oracle_chain(orc, f, [0.0, 0.0, h], beta, 2500, 5, rng)
        st = estimators(E, M, beta, N)
        st["absQ"] = np.abs(Q).mean()
        ref.append(st)
    # the plaquette charge on a torus is an integer for unit spins; the stored
    # fp32 spins are unit vectors to ~1e-7, which moves a triangle's solid
    # angle by ~1e-7 / |s_i + s_j|, so a sample holding a nearly antiparallel
    # neighbour pair (|s_i + s_j| ~ 1e-3, about one per 10^5 samples here)
    # is off by more than 1e-5.  (The charge of given spins matches the
    # oracle's FP64 calc_TC to 1e-10: test_measure_matches_oracle.)
    off = np.abs(qw[:, :, 0] - np.round(qw[:, :, 0]))
    print("charge: %d of %d samples off an integer by > 1e-5, max %.3g" % ((off > 1e-5).sum(), off.size, off.max()))
    assert (off > 1e-5).sum() <= 1e-4 * off.size and off.max() < 1e-2
    for k in ("E", "absM", "C", "chi", "absQ"):
        m, e = replica_stats([p[k] for p in dev])
        mr, er = replica_stats([p[k] for p in ref])
        sig = math.hypot(e, er)
        print("%-5s device %.5f +- %.5f   oracle %.5f +- %.5f   (%.2f sigma)" % (k, m, e, mr, er, abs(m - mr) / sig))
        assert abs(m - mr) < 3 * sig, (k, m, e, mr, er)
    assert np.mean([p["absQ"] for p in ref]) > 0.05     # the charge does fluctuate at this T


def test_heisenberg_3d_cone_observables_match_oracle_chain(orc, bloch3d):
    # the benchmark's move: cone trials whose half-angles were adapted on the
    # device (nmc_sweep_adapt) and then frozen, on a 3-D periodic lattice
    # with Bloch DMI and a field; the oracle's chain uses the reference's
    # uniform-sphere trials.  Both must sample the same distribution.
    from nanomagmc_amd import Bond, Lattices
    L, h, T = 4, 0.2, 1.2
    beta = 1.0 / T
    d, Jv, Dv = orc.parse_bonds(bloch3d, 3, 1.0, 0.3)
    R, ns = 16, 6000
    lat = Lattices(3, False, True, [Bond(d[k], Jv[k], Dv[k]) for k in range(len(Jv))],
                   [{"n": (L, L, L), "edge": L}] * R, use_d=True, trial="cone", seed=0x4E4D4314)
    lat.set_params([beta] * R, h)
    lat.set_cone([1.0] * R)
    lat.randomize()
    th = lat.equil_adapt(400, 10)
    assert np.all(th < 3.0)        # adapted away from the whole sphere at this T
    lat.equil(100)
    lat.series_reset(ns, with_charge=False)
    lat.run_samples(ns, 5)
    obs, _, _ = lat.get_series()
    N = L ** 3
    dev = [estimators(obs[i, :, 10], obs[i, :, 0:3], beta, N) for i in range(R)]
    f = orc.Field(box_locs(3, L), 3, L, 1, d, Jv, Dv)
    rng = np.random.default_rng(0x4E4D4315)
    ref = []
    for _ in range(8):
        E, M, _ = oracle_chain(orc, f, [0.0, 0.0, h], beta, 2000, 5, rng)
        ref.append(estimators(E, M, beta, N))
    for k in ("E", "absM", "C", "chi"):
        m, e = replica_stats([p[k] for p in dev])
        mr, er = replica_stats([p[k] for p in ref])
        sig = math.hypot(e, er)
        print("%-5s device %.5f +- %.5f   oracle %.5f +- %.5f   (%.2f sigma)" % (k, m, e, mr, er, abs(m - mr) / sig))
        assert abs(m - mr) < 3 * sig, (k, m, e, mr, er)
