"""Pins the oracle (the CPU checker) against the reference's own known
answers (Old/tests/*.hpp), closed forms and brute force.  CPU only."""
import math

import numpy as np
import pytest

from tests import exact
from tests.fixtures import afm_spins, box_locs, neel2d, nn2d, skyrm_spins, up_spins


def field(orc, dim, edge, periodic, bd, J, D, locs=None):
    if locs is None:
        locs = box_locs(dim, edge)
    return orc.Field(locs, dim, edge, periodic, bd, J, D)


# ---------------------------------------------------------------- Philox ----
def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32 10 rounds
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, out in kat:
        assert list(orc.philox(ctr, key)) == out


def test_philox_matches_rocrand(orc):
    rng = np.random.default_rng(1)
    for _ in range(500):
        ctr = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        key = rng.integers(0, 2 ** 32, 2, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(orc.philox(ctr, key), orc.rocrand_philox(ctr, key))


# ---------------------------------------------------------- parser/builder --
def test_parse_bloch_file(orc, bloch3d):
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.5)
    assert d.tolist() == [[0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0]]
    assert np.all(J == 1.0)
    assert np.allclose(D, 0.5 * d)          # D parallel to r-hat, scaled by DMISTREN


def test_parse_scaling_and_2d_misparse(orc, bloch3d):
    # a 3-D file read as 2-D shifts every column (hazard documented for the
    # reference parser): 6 records of 7 numbers become 7 records of 6
    d, J, D = orc.parse_bonds(bloch3d, 2, 2.0, 1.0)
    assert len(J) == 7
    assert d[0].tolist() == [0, 0, 0] and J[0] == 2.0 * 1


def test_build_square_cube_counts(orc):
    assert len(orc.build_sites(2, 10, "s")) == 100
    locs = orc.build_sites(3, 6, "c")
    assert len(locs) == 216
    assert np.array_equal(locs, box_locs(3, 6))


def test_build_faithful_carry_bug(orc):
    edge = 5
    faithful = orc.build_sites(3, edge, "c", faithful_carry=True)
    assert len(faithful) == edge ** 3
    bad = faithful[faithful[:, 1] == edge]
    assert sorted(bad[:, 0].tolist()) == list(range(edge - 1))       # (x, edge, 0), x < edge-1
    assert np.all(bad[:, 2] == 0)
    present = {tuple(p) for p in faithful.tolist()}
    missing = [x for x in range(1, edge) if (x, 0, 0) not in present]
    assert missing == list(range(1, edge))
    # 2-D has no carry bug
    assert np.array_equal(orc.build_sites(2, 7, "s", faithful_carry=True), box_locs(2, 7))


def test_weibull_keep_probability(orc):
    rng = np.random.default_rng(3)
    edge, a, wb = 40, 10.0, 8.0
    u = rng.random(edge * edge)
    locs = orc.build_sites(2, edge, "w", a, wb, u)
    r0 = orc.weibull_r0(wb)
    assert abs(r0 - 1 / math.gamma(1 + 1 / wb)) < 1e-14
    # expected kept count from the closed form keep probability
    g = box_locs(2, edge)[:, :2].astype(float)
    c = (edge - 1) / 2
    dist = np.sqrt((((g - c) / a) ** 2).sum(axis=1))
    p = np.exp(-((dist / r0) ** wb))
    mean, var = p.sum(), (p * (1 - p)).sum()
    assert abs(len(locs) - mean) < 5 * math.sqrt(var)
    # and site-by-site: kept iff u < p (same uniforms)
    kept = set(map(tuple, locs[:, :2].tolist()))
    want = {tuple(map(int, g[i])) for i in range(len(g)) if u[i] < p[i]}
    assert kept == want


def _literal_set_neigh(locs, dim, edge, periodic, bd):
    """Old/lib/field_type.cpp:112-141 written out as the O(N^2) loops."""
    n = len(locs)
    neigh = []
    for i in range(n):
        row = []
        for k in range(len(bd)):
            t = locs[i] + bd[k]
            if periodic:
                t = np.array([(edge + (int(t[a]) % edge if t[a] >= 0 else -((-int(t[a])) % edge))) % edge
                              if a < dim else t[a] for a in range(3)])
            for j in range(n):
                if i == j:
                    continue
                if np.array_equal(t, locs[j]):
                    row.append((j, k))
                    break
        neigh.append(row)
    return neigh


@pytest.mark.parametrize("dim,edge,periodic", [(3, 4, 1), (3, 4, 0), (2, 5, 0), (2, 4, 1), (2, 2, 1)])
def test_set_neigh_matches_literal_search(orc, dim, edge, periodic):
    bd, J, D = orc.nn_bonds(dim)
    rng = np.random.default_rng(dim * 10 + edge)
    locs = box_locs(dim, edge)
    if not periodic:
        locs = locs[rng.random(len(locs)) < 0.8]
    f = field(orc, dim, edge, periodic, bd, J, D, locs)
    lit = _literal_set_neigh(locs, dim, edge, periodic, bd)
    for i in range(len(locs)):
        got = [(int(f.neigh[i, q]), int(f.nchoice[i, q])) for q in range(f.ncount[i])]
        assert got == lit[i]


def test_neighbour_counts(orc):
    L = 6
    bd, J, D = orc.nn_bonds(3)
    assert field(orc, 3, L, 0, bd, J, D).ncount.sum() == 2 * 3 * L * L * (L - 1)
    assert field(orc, 3, L, 1, bd, J, D).ncount.sum() == 6 * L ** 3


# ------------------------------------------------- reference known answers --
@pytest.fixture(scope="module")
def ref_fields(orc, bloch3d):
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.0)
    open_ = field(orc, 3, 10, 0, d, J, D)
    per = field(orc, 3, 10, 1, d, J, D)
    return open_, per


def test_ising_3d_known_answers(ref_fields):
    # Old/tests/ising_test.hpp:122-164
    f, fp = ref_fields
    fm = up_spins(1000, True)
    afm = afm_spins(f.locs, 3, True)
    zero, hx = [0, 0, 0], [0.1, 0, 0]
    assert f.energy(fm, zero) == -2700 and fp.energy(fm, zero) == -3000
    assert f.energy(fm, hx) == -2800 and fp.energy(fm, hx) == -3100
    assert f.mag(fm)[0] == 1000 and fp.mag(fm)[0] == 1000
    assert f.submag(fm)[0] == 500 and fp.submag(fm)[0] == 500
    assert f.energy(afm, zero) == 2700 and f.energy(afm, hx) == 2700
    assert f.mag(afm)[0] == 0 and f.submag(afm)[0] == 500


def test_heisenberg_3d_known_answers(ref_fields):
    # Old/tests/heis_test.hpp:285-342
    f, _ = ref_fields
    fm = up_spins(1000, False)
    afm = afm_spins(f.locs, 3, False)
    hz = [0, 0, 0.1]
    assert f.energy(fm, [0, 0, 0]) == -2700 and f.energy(afm, [0, 0, 0]) == 2700
    assert f.energy(fm, hz) == -2800 and f.energy(afm, hz) == 2700
    assert f.mag(fm).tolist() == [0, 0, 1000] and f.mag(afm).tolist() == [0, 0, 0]
    assert f.submag(fm).tolist() == [0, 0, 500] and f.submag(afm).tolist() == [0, 0, 500]
    assert f.sub4mag(fm).tolist() == [0, 0, 63] and f.sub4mag(afm).tolist() == [0, 0, 63]


def test_2d_closed_forms(orc):
    # the commented 2-D gtests (ising_test.hpp:17-97): 10x10 open, 180 bonds
    d, J, D = nn2d()
    f = field(orc, 2, 10, 0, d, J, D)
    fm = up_spins(100, True)
    assert f.energy(fm, [0, 0, 0]) == -180
    assert f.energy(fm, [0.1, 0, 0]) == -190
    assert f.energy(afm_spins(f.locs, 2, True), [0, 0, 0]) == 180
    assert f.mag(fm)[0] == 100 and f.submag(fm)[0] == 50
    bulk = 5 * 10 + 5
    assert f.delta_e(fm, bulk, [-1, 0, 0], [0, 0, 0]) == 8


@pytest.mark.parametrize("ising,use_d", [(True, False), (False, False), (False, True)])
def test_delta_e_consistent_with_sampler_energy(orc, bloch3d, ising, use_d):
    # ising_test.hpp:166-181 / heis_test.hpp:344-376, extended to DMI: dE_func
    # is the difference of E_s (convention 1) for any bond table with
    # D_ji = -D_ij
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.75)
    f = field(orc, 3, 6, 1, d, J, D)
    rng = np.random.default_rng(7)
    from tests.fixtures import random_unit
    s = np.sign(rng.random((f.n, 1)) - 0.5) * np.array([[1, 0, 0]]) if ising else random_unit(f.n, rng)
    H = [0.1, 0, 0] if ising else [0.02, -0.05, 0.1]
    E = f.energy(s, H, use_d=use_d, convention=1)
    for _ in range(300):
        i = int(rng.integers(f.n))
        t = np.array([rng.choice([-1.0, 1.0]), 0, 0]) if ising else random_unit(1, rng)[0]
        dE = f.delta_e(s, i, t, H, use_d=use_d)
        s[i] = t
        E2 = f.energy(s, H, use_d=use_d, convention=1)
        assert abs(E + dE - E2) <= 1e-10 * max(1.0, abs(E2))
        E = E2


def test_reported_energy_disagrees_with_sampler_under_dmi(orc, bloch3d):
    # E_func's D x s_j term is the opposite orientation of dE_func's s_j x D
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.75)
    f = field(orc, 3, 6, 1, d, J, D)
    rng = np.random.default_rng(8)
    from tests.fixtures import random_unit
    s = random_unit(f.n, rng)
    E_r, E_s = f.energy(s, [0, 0, 0], use_d=True, convention=0), f.energy(s, [0, 0, 0], use_d=True, convention=1)
    assert abs(E_r - E_s) > 1e-3
    # and the two agree without DMI
    assert abs(f.energy(s, [0, 0, 0], convention=0) - f.energy(s, [0, 0, 0], convention=1)) < 1e-9


def test_dmi_dimer_closed_form(orc):
    # E_s of a single 2-D Neel bond: -J s1.s2 - 2 * (1/2) D_12 . (s1 x s2)
    d, J, D = neel2d(0.3)
    locs = np.array([[0, 0, 0], [1, 0, 0]], dtype=np.int32)
    f = orc.Field(locs, 2, 2, 0, d, J, D)
    s1, s2 = np.array([0.0, 0.6, 0.8]), np.array([0.48, -0.64, 0.6])
    s = np.stack([s1, s2])
    D12 = 0.3 * np.cross([0, 0, 1.0], [1.0, 0, 0])
    want = -np.dot(s1, s2) - np.dot(D12, np.cross(s1, s2))
    assert abs(f.energy(s, [0, 0, 0], use_d=True, convention=1) - want) < 1e-14


# ---------------------------------------------------------------- charges --
def test_solid_angle_closed_form(orc):
    x, y, z = np.eye(3)
    assert abs(orc.solid_angle(x, y, z) - math.pi / 4) < 1e-15   # half of the octant's pi/2
    assert abs(orc.solid_angle(x, z, y) + math.pi / 4) < 1e-15   # orientation flips the sign
    assert orc.solid_angle(z, z, z) == 0.0


def test_tcharge_reference_fixtures(orc, bloch3d):
    # Old/tests/heis_test.hpp:378-393
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.75)
    f = field(orc, 3, 10, 0, d, J, D)
    assert np.all(f.tcharge(up_spins(1000, False)) == 0.0)
    locs, s = skyrm_spins(30, 3)
    fs = orc.Field(locs, 3, 30, 0, d, J, D)
    q = fs.tcharge(s)
    assert q.shape == (30,)
    assert np.all(np.abs(q + 1) < 0.01)


def test_tcharge_wrap_integer_on_torus(orc):
    # a periodic 2-D skyrmion: the wrapped charge is an integer, the
    # reference's (no wrap) is not counted across the boundary
    L = 24
    d, J, D = nn2d()
    locs = box_locs(2, L)
    f = orc.Field(locs, 2, L, 1, d, J, D)
    rng = np.random.default_rng(5)
    from tests.fixtures import random_unit
    s = random_unit(f.n, rng)
    qw = f.tcharge(s, wrap=True)[0]
    assert abs(qw - round(qw)) < 1e-9
    assert f.tcharge(s, wrap=False)[0] != pytest.approx(qw, abs=1e-6)


# ---------------------------------------------------------------- misc -----
def test_atoln_closed_form(orc):
    m, s = orc.atoln(10.0, 2.5)
    assert abs(m - math.log(100 / math.sqrt(106.25))) < 1e-14
    assert abs(s - math.sqrt(math.log(1.0625))) < 1e-14


def test_accept_threshold_brute(orc):
    rng = np.random.default_rng(11)
    for _ in range(200):
        dE = float(rng.choice([4.0, 8.0, 0.2, 12.0, 3.7, 1e-9]))
        beta = float(rng.uniform(0.1, 3.0))
        T = orc.accept_threshold(dE, beta)
        pred = lambda k: (-dE * beta) > (math.log(k / 4294967296.0) if k else -math.inf)
        assert pred(T)
        if T < 2 ** 32 - 1:
            assert not pred(T + 1)
    assert orc.accept_threshold(-1.0, 1.0) == 2 ** 32 - 1
    assert orc.accept_threshold(0.0, 1.0) == 2 ** 32 - 1


# ------------------------------------------------------ exact enumeration --
def test_kaufman_matches_enumeration():
    E, M = exact.ising_periodic_energies(4, 4)
    for T in (1.5, 2.269, 3.0):
        e_enum = exact.enum_averages(E, M, 1 / T)["E"]
        assert abs(exact.kaufman_energy(4, T) - e_enum) < 1e-6


def test_kaufman_tends_to_onsager():
    for T in (1.5, 3.5):
        assert abs(exact.kaufman_energy(64, T) / 64 ** 2 - exact.onsager_energy_per_site(T)) < 1e-6


def test_random_site_chain_matches_enumeration(orc):
    # state::equil with fed uniforms on a 4x4 periodic Ising lattice
    d, J, D = nn2d()
    f = field(orc, 2, 4, 1, d, J, D)
    E, M = exact.ising_periodic_energies(4, 4)
    T = 2.5
    want = exact.enum_averages(E, M, 1 / T)["E"]
    rng = np.random.default_rng(21)
    s = up_spins(16, True)
    samples = []
    for _ in range(40):
        n = 100000
        sites = rng.integers(0, 16, n)
        trials = np.zeros((n, 3))
        trials[:, 0] = rng.integers(0, 2, n) * 2 - 1
        u = rng.integers(0, 2 ** 32, n) / 4294967296.0
        f.equil_replay(s, [0, 0, 0], 1 / T, sites, trials, u)
        samples.append(f.energy(s, [0, 0, 0]))
        for _ in range(50):
            f.equil_replay(s, [0, 0, 0], 1 / T, sites[:160], trials[:160], u[:160])
            samples.append(f.energy(s, [0, 0, 0]))
    m, err, _ = exact.mean_err(samples)
    assert abs(m - want) < 4 * err + 1e-3


def test_checkerboard_schedule_matches_enumeration(orc):
    # the device's visit order and Philox stream, restated on the CPU
    d, J, D = nn2d()
    f = field(orc, 2, 4, 1, d, J, D)
    E, M = exact.ising_periodic_energies(4, 4)
    e, c = f.colour_index((4, 4))
    for T in (2.0, 3.0):
        want = exact.enum_averages(E, M, 1 / T)
        s = up_spins(16, True)
        es, ms = [], []
        for sw in range(20000):
            f.checkerboard_ising(s, [0, 0, 0], 1 / T, e, c, seed=99, inst=0, sweep0=sw, nsweeps=1)
            es.append(f.energy(s, [0, 0, 0]))
            ms.append(abs(s[:, 0].sum()))
        m, err, _ = exact.mean_err(es)
        assert abs(m - want["E"]) < 4 * err
        m, err, _ = exact.mean_err(ms)
        assert abs(m - want["absM"]) < 4 * err
