"""The device against the committed golden vectors (tests/golden/vectors.npz,
the oracle's output, pinned on the CPU by tests/test_golden.py).  Needs an
MI355X; reads no oracle at run time.

* dE through nmc_delta_e: Ising bit for bit; Heisenberg within 1e-5 of the
  site's own scale |H + V| |s - t| (fp32 local field);
* the fed Ising replay through nmc_replay_ising: every decision equal;
* the 30^3 skyrmion's charge per slice through nmc_measure.
"""
import numpy as np
import pytest

from tests.fixtures import skyrm_spins
from tests.golden import make_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gold():
    with np.load(make_golden.OUT, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _lattice(dim, edge, periodic, ising, use_d, dirs, J, D):
    from nanomagmc_amd import Bond, Lattices
    bonds = [Bond(tuple(int(x) for x in dirs[k][:dim]), float(J[k]), tuple(float(x) for x in D[k]))
             for k in range(len(J))]
    return Lattices(dim, ising, bool(periodic), bonds, [{"n": (edge,) * dim, "edge": edge}], use_j=True,
                    use_d=use_d)


@pytest.mark.parametrize("name", sorted(make_golden.DE_CASES))
def test_delta_e_against_golden(gold, name):
    dim, edge, periodic, ising, use_d, _ = make_golden.DE_CASES[name]
    p = "de_%s_" % name
    lat = _lattice(dim, edge, periodic, ising, use_d, gold[p + "dirs"], gold[p + "J"], gold[p + "D"])
    lat.set_params(0.7, float(gold[p + "h"]))
    s = gold[p + "spins"]
    lat.set_spins(0, s)
    got = lat.delta_e(0, gold[p + "sites"], gold[p + "trials"])
    want = gold[p + "want"]
    if ising:
        assert np.array_equal(got, want)
    else:
        ds = np.linalg.norm(s[gold[p + "sites"]] - gold[p + "trials"], axis=1)
        tol = 1e-5 * gold[p + "scale"] + 1e-6 * ds
        assert np.all(np.abs(got - want) <= tol), np.max(np.abs(got - want) - tol)


def test_replay_against_golden(gold):
    p = "rp_ising2d_"
    dirs, J, D = make_golden.nn2d()
    lat = _lattice(2, 16, 1, True, False, dirs, J, D)
    lat.set_params(1.0 / float(gold[p + "T"]), float(gold[p + "h"]))
    lat.set_spins(0, gold[p + "spins"])
    got = lat.replay_ising(0, gold[p + "sites"], gold[p + "trials"], gold[p + "k"])
    assert np.array_equal(np.asarray(got, dtype=np.uint8), gold[p + "accept"])


def test_skyrmion_charge_against_golden(gold, orc):
    d, J, D = orc.parse_bonds(make_golden.BLOCH, 3, 1.0, 0.3)
    lat = _lattice(3, 30, 0, False, True, d, J, D)
    _, s = skyrm_spins(30, 3)
    lat.set_spins(0, s.astype(np.float32).astype(np.float64))
    _, qr, _ = lat.measure()
    assert np.allclose(qr[0], gold["tc_skyrm30"], rtol=0, atol=1e-10)
