"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's known answers.  Needs an MI355X."""
import math

import numpy as np
import pytest

from tests import exact
from tests.fixtures import afm_spins, box_locs, neel2d, nn2d, random_unit, skyrm_spins, up_spins
from tests.gpu_util import make_pair

pytestmark = pytest.mark.gpu


def f32(s):
    return np.asarray(s, dtype=np.float32).astype(np.float64)


# ---------------------------------------------------- reference known answers
@pytest.mark.parametrize("ising", [True, False])
def test_known_answers_3d(orc, bloch3d, ising):
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.0)
    for periodic in (0, 1):
        lat, f = make_pair(orc, 3, 10, periodic, d, J, D, ising)
        H = 0.1
        for spins, want_E0, want_Eh in (
                (up_spins(1000, ising), -3000 if periodic else -2700, -3100 if periodic else -2800),
                (afm_spins(f.locs, 3, ising), None, None)):
            lat.set_spins(0, spins)
            for h, want in ((0.0, want_E0), (H, want_Eh)):
                lat.set_params(1.0, h)
                obs, _, _ = lat.measure()
                if want is not None:
                    assert obs[0, 9] == want
                Hv = [h, 0, 0] if ising else [0, 0, h]
                assert obs[0, 9] == f.energy(spins, Hv)
                a = 0 if ising else 2
                assert obs[0, a] == f.mag(spins)[a]
                assert obs[0, 3 + a] == f.submag(spins)[a]
                assert obs[0, 6 + a] == f.sub4mag(spins)[a]
        if not periodic:
            lat.set_spins(0, afm_spins(f.locs, 3, ising))
            lat.set_params(1.0, 0.1)
            obs, _, _ = lat.measure()
            assert obs[0, 9] == 2700 and obs[0, 3 + (0 if ising else 2)] == 500
            if not ising:
                assert obs[0, 8] == 63


# ------------------------------------------------- random configurations ----
CASES = [
    # dim, edge, periodic, ising, use_d, masked
    (2, 16, 1, True, False, False),
    (2, 15, 0, True, True, True),
    (3, 8, 1, True, False, False),
    (3, 7, 0, True, False, True),
    (2, 16, 1, False, True, False),
    (2, 13, 0, False, True, True),
    (3, 8, 1, False, True, False),
    (3, 9, 0, False, True, True),
    (2, 16, 1, True, False, True),
    (2, 14, 1, False, True, True),
    (3, 8, 1, False, True, True),
    (3, 18, 0, False, True, True),   # several measurement tiles per colour
    (3, 16, 1, False, False, True),
]


def _case(orc, bloch3d, dim, edge, periodic, ising, use_d, masked, rng, trial="sphere"):
    if dim == 3:
        bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.3)
    else:
        bd, J, D = neel2d(0.3)
    locs = None
    if masked:
        locs = box_locs(dim, edge)
        locs = locs[rng.random(len(locs)) < 0.75]
    lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, ising, use_d=use_d, locs=locs, trial=trial)
    if ising:
        s = np.zeros((f.n, 3))
        s[:, 0] = rng.choice([-1.0, 1.0], f.n)
    else:
        s = f32(random_unit(f.n, rng))
    lat.set_spins(0, s)
    assert np.array_equal(lat.get_spins(0), s)
    return lat, f, s


@pytest.mark.parametrize("case", CASES)
def test_measure_matches_oracle(orc, bloch3d, case):
    dim, edge, periodic, ising, use_d, masked = case
    rng = np.random.default_rng(hash(case) % 2 ** 32)
    lat, f, s = _case(orc, bloch3d, *case, rng)
    h = 0.1
    lat.set_params(0.7, h)
    obs, qr, qw = lat.measure()
    H = [h, 0, 0] if ising else [0, 0, h]
    want = np.concatenate([f.mag(s), f.submag(s), f.sub4mag(s),
                           [f.energy(s, H, use_d=use_d, convention=0), f.energy(s, H, use_d=use_d, convention=1)]])
    if ising:
        assert np.array_equal(obs[0], want)       # integer-valued: bit-exact
    else:
        assert np.allclose(obs[0], want, rtol=1e-12, atol=1e-9)
        assert np.allclose(qr[0], f.tcharge(s, wrap=False), rtol=1e-10, atol=1e-10)
        assert np.allclose(qw[0], f.tcharge(s, wrap=True), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("dim,periodic,masked", [(2, 1, False), (2, 0, True), (3, 1, False), (3, 0, True),
                                                 (3, 1, True)])
def test_measure_general_dmi_matches_oracle(orc, bloch3d, dim, periodic, masked):
    # DMI vectors off the lattice's one-axis pattern: the measurement's
    # six-term form of n x D
    rng = np.random.default_rng(401 + 10 * dim + periodic)
    edge = 8 if dim == 3 else 14
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.3) if dim == 3 else neel2d(0.3)
    D = np.asarray(D) + 0.1 * np.sign(np.asarray(D).sum(axis=1, keepdims=True))
    locs = None
    if masked:
        locs = box_locs(dim, edge)
        locs = locs[rng.random(len(locs)) < 0.75]
    lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, False, use_d=True, locs=locs)
    s = f32(random_unit(f.n, rng))
    lat.set_spins(0, s)
    lat.set_params(0.7, 0.1)
    obs, qr, qw = lat.measure()
    want = np.concatenate([f.mag(s), f.submag(s), f.sub4mag(s),
                           [f.energy(s, [0, 0, 0.1], use_d=True, convention=0),
                            f.energy(s, [0, 0, 0.1], use_d=True, convention=1)]])
    assert np.allclose(obs[0], want, rtol=1e-12, atol=1e-9)
    assert abs(obs[0, 9] - obs[0, 10]) > 1e-3          # the DMI term is present
    assert np.allclose(qr[0], f.tcharge(s, wrap=False), rtol=1e-10, atol=1e-10)
    assert np.allclose(qw[0], f.tcharge(s, wrap=True), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("dim,periodic,masked", [(2, 1, False), (2, 0, True), (3, 1, False), (3, 0, True)])
def test_measure_asymmetric_table_matches_oracle(orc, bloch3d, dim, periodic, masked):
    # a bond and its reverse with different J and D (the interaction file
    # lists each direction on its own): the per-site form of the local-field
    # sums; symmetric tables take the per-bond form, tested everywhere else
    rng = np.random.default_rng(503 + 10 * dim + periodic)
    edge = 8 if dim == 3 else 14
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.3) if dim == 3 else neel2d(0.3)
    J = np.asarray(J, dtype=np.float64) * (1.0 + 0.3 * np.arange(len(J)))
    D = np.asarray(D) * (1.0 + 0.2 * np.arange(len(J)))[:, None]
    locs = None
    if masked:
        locs = box_locs(dim, edge)
        locs = locs[rng.random(len(locs)) < 0.75]
    lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, False, use_d=True, locs=locs)
    s = f32(random_unit(f.n, rng))
    lat.set_spins(0, s)
    lat.set_params(0.7, 0.1)
    obs, _, _ = lat.measure()
    want = np.concatenate([f.mag(s), f.submag(s), f.sub4mag(s),
                           [f.energy(s, [0, 0, 0.1], use_d=True, convention=0),
                            f.energy(s, [0, 0, 0.1], use_d=True, convention=1)]])
    assert np.allclose(obs[0], want, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("dim,periodic,masked", [(3, 1, False), (2, 1, False), (3, 0, True)])
def test_measure_without_charges_is_identical(orc, bloch3d, dim, periodic, masked):
    # a measurement that asks for no charges skips the triangles: every other
    # observable must be bitwise the one of the full measurement, directly and
    # in a sample series
    import ctypes as C
    from nanomagmc_amd._lib import check, lib
    rng = np.random.default_rng(77 + dim)
    edge = 8 if dim == 3 else 16
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.3) if dim == 3 else neel2d(0.3)
    locs = None
    if masked:
        locs = box_locs(dim, edge)
        locs = locs[rng.random(len(locs)) < 0.7]
    lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, False, use_d=True, locs=locs, ninst=3)
    lat.set_params([0.5, 1.0, 2.0], [0.0, 0.05, 0.1])
    lat.randomize()
    lat.equil(5)
    full, _, _ = lat.measure()
    bare = np.zeros_like(full)
    check(lib().nmc_measure(lat._h, None, 0, bare.ctypes.data_as(C.POINTER(C.c_double)), None, None))
    assert np.array_equal(full, bare)
    series = []
    for with_charge in (True, False):
        lat.series_reset(2, with_charge=with_charge)
        lat.sample(0)
        lat.sample(1)
        series.append(lat.get_series()[0])
    assert np.array_equal(series[0], series[1])
    assert np.array_equal(series[0][:, 0], full)


def test_tcharge_skyrmion_fixture(orc, bloch3d):
    # Old/tests/heis_test.hpp:378-393 on the device
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.75)
    locs, s = skyrm_spins(30, 3)
    lat, f = make_pair(orc, 3, 30, 0, d, J, D, False)
    s = f32(s)
    lat.set_spins(0, s)
    _, qr, _ = lat.measure()
    assert qr.shape == (1, 30)
    assert np.all(np.abs(qr[0] + 1) < 0.01)
    assert np.allclose(qr[0], f.tcharge(s), atol=1e-12)
    lat.set_spins(0, up_spins(27000, False))
    _, qr, _ = lat.measure()
    assert np.all(qr[0] == 0)


# ------------------------------------------------------------------ dE ------
@pytest.mark.parametrize("case", CASES)
def test_delta_e_matches_oracle(orc, bloch3d, case):
    dim, edge, periodic, ising, use_d, masked = case
    rng = np.random.default_rng(1 + hash(case) % 2 ** 31)
    lat, f, s = _case(orc, bloch3d, *case, rng)
    h = 0.1
    lat.set_params(0.7, h)
    H = [h, 0, 0] if ising else [0, 0, h]
    n = 256
    sites = rng.integers(0, f.n, n)
    if ising:
        trials = np.zeros((n, 3))
        trials[:, 0] = rng.choice([-1.0, 1.0], n)
    else:
        trials = f32(random_unit(n, rng))
    got = lat.delta_e(0, sites, trials)
    want = np.array([f.delta_e(s, sites[i], trials[i], H, use_d=use_d) for i in range(n)])
    if ising:
        assert np.array_equal(got, want)          # bit-exact
    else:
        # fp32 local field: within 1e-5 of the site's own scale |H + V| |s - t|,
        # V = sum_j J_j s_j + s_j x D_j from the oracle's neighbour lists (the
        # field dE_func dots with s - t, thermodynamics.cpp:51-69); the second
        # term is fp32 rounding of the ~z terms of V where they cancel
        for i in range(n):
            j = sites[i]
            V = np.array([0.0, 0.0, h])
            terms = abs(h)
            for c in range(f.ncount[j]):
                k, nb = f.nchoice[j, c], s[f.neigh[j, c]]
                V += f.J[k] * nb + (np.cross(nb, f.D[k]) if use_d else 0.0)
                terms += abs(f.J[k]) + (np.linalg.norm(f.D[k]) if use_d else 0.0)
            ds = np.linalg.norm(s[j] - trials[i])
            tol = 1e-5 * np.linalg.norm(V) * ds + 4 * 2.0 ** -24 * terms * ds
            assert abs(got[i] - want[i]) <= tol, (i, got[i], want[i], tol)


# ------------------------------------------------ accept decisions (Ising) ---
@pytest.mark.parametrize("dim,edge,periodic", [(2, 16, 1), (3, 6, 0), (2, 9, 0)])
def test_replay_accept_decisions_identical(orc, bloch3d, dim, edge, periodic):
    rng = np.random.default_rng(dim * 100 + edge)
    bd, J, D = (orc.parse_bonds(bloch3d, 3, 1.0, 0.0) if dim == 3 else nn2d())
    lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, True)
    s = np.zeros((f.n, 3))
    s[:, 0] = rng.choice([-1.0, 1.0], f.n)
    lat.set_spins(0, s)
    T, h = 2.3, 0.1
    lat.set_params(1 / T, h)
    n = 20000
    sites = rng.integers(0, f.n, n).astype(np.int32)
    tr = rng.choice([-1, 1], n).astype(np.int8)
    k = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    got = lat.replay_ising(0, sites, tr, k)
    trials = np.zeros((n, 3))
    trials[:, 0] = tr
    acc = f.equil_replay(s, [h, 0, 0], 1 / T, sites, trials, k / 4294967296.0)
    assert np.array_equal(got, acc)
    assert np.array_equal(lat.get_spins(0), s)


def test_replay_boundary_uniforms(orc):
    # uniforms placed on k = T and T + 1 for each flip's threshold
    d, J, D = nn2d()
    lat, f = make_pair(orc, 2, 8, 1, d, J, D, True)
    rng = np.random.default_rng(4)
    T, h = 1.7, 0.05
    lat.set_params(1 / T, h)
    s = up_spins(64, True)
    lat.set_spins(0, s)
    sites, trials, ks = [], [], []
    cur = s.copy()
    for _ in range(4000):
        i = int(rng.integers(64))
        t = -cur[i, 0]
        dE = f.delta_e(cur, i, [t, 0, 0], [h, 0, 0])
        thr = orc.accept_threshold(dE, 1 / T)
        k = min(thr + int(rng.integers(0, 2)), 2 ** 32 - 1)
        acc = dE <= 0 or ((-dE / T) > (math.log(k / 4294967296.0) if k else -math.inf))
        if acc:
            cur[i, 0] = t
        sites.append(i); trials.append(int(t)); ks.append(k)
    got = lat.replay_ising(0, np.array(sites), np.array(trials, dtype=np.int8), np.array(ks, dtype=np.uint32))
    tr = np.zeros((len(sites), 3))
    tr[:, 0] = trials
    want = f.equil_replay(s, [h, 0, 0], 1 / T, np.array(sites), tr, np.array(ks) / 4294967296.0)
    assert np.array_equal(got, want)
    assert 0 < want.sum() < len(want)


# ------------------------------------ whole checkerboard sweeps (Ising) ------
@pytest.mark.parametrize("dim,edge,periodic,masked", [(2, 16, 1, False), (2, 11, 0, True), (3, 8, 1, False),
                                                      (3, 7, 0, True), (2, 12, 1, True), (3, 6, 1, True)])
def test_checkerboard_trajectory_identical(orc, bloch3d, dim, edge, periodic, masked):
    rng = np.random.default_rng(edge + 17 * dim)
    bd, J, D = (orc.parse_bonds(bloch3d, 3, 1.0, 0.0) if dim == 3 else nn2d())
    locs = None
    if masked:
        locs = box_locs(dim, edge)
        locs = locs[rng.random(len(locs)) < 0.8]
    seed = 0x4E4D4301
    lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, True, locs=locs, seed=seed, ninst=3)
    T, h = 2.4, 0.07
    lat.set_params([1 / T, 1 / 2.0, 1 / 3.0], [h, 0.0, -0.1])
    params = [(1 / T, h), (1 / 2.0, 0.0), (1 / 3.0, -0.1)]
    e, c = f.colour_index((edge,) * dim)
    starts = []
    for i in range(3):
        s = np.zeros((f.n, 3))
        s[:, 0] = rng.choice([-1.0, 1.0], f.n)
        lat.set_spins(i, s)
        starts.append(s)
    lat.equil(4)
    lat.equil(3, inst=[1, 2])
    for i in range(3):
        beta, hh = params[i]
        s = starts[i]
        nsw = 4 if i == 0 else 7
        f.checkerboard_ising(s, [hh, 0, 0], beta, e, c, seed=seed, inst=i, sweep0=0, nsweeps=nsw)
        assert np.array_equal(lat.get_spins(i), s), "instance %d" % i


@pytest.mark.parametrize("ising,dim,periodic,masked,trial,general", [
    (True, 2, 1, False, "sphere", False), (True, 3, 0, True, "sphere", False),
    (False, 2, 1, False, "sphere", False), (False, 3, 0, True, "cone", False), (False, 3, 1, False, "cone", False),
    (False, 2, 0, True, "sphere", False), (False, 3, 0, True, "sphere", True), (False, 2, 1, False, "cone", True),
    (True, 2, 1, True, "sphere", False), (False, 3, 1, True, "cone", False), (False, 2, 1, True, "sphere", True)])
def test_lds_resident_and_streaming_paths_agree(orc, bloch3d, ising, dim, periodic, masked, trial, general):
    # the LDS-resident launch (whole sweeps in one workgroup) and the
    # streaming launches (one per colour pass) draw the same numbers; with
    # `general` the DMI vectors are not on one spin axis per bond (the
    # kernels' six-term form instead of the two-term one)
    rng = np.random.default_rng(31 + dim)
    edge = 10 if dim == 3 else 24
    bd, J, D = (orc.parse_bonds(bloch3d, 3, 1.0, 0.2) if dim == 3 else neel2d(0.2))
    if general:
        D = np.asarray(D) + 0.1 * np.sign(np.asarray(D).sum(axis=1, keepdims=True))
    locs = None
    if masked:
        locs = box_locs(dim, edge)
        locs = locs[rng.random(len(locs)) < 0.8]
    lats = []
    for path in ("auto", "stream"):
        lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, ising, use_d=not ising, locs=locs, trial=trial,
                           ninst=3, seed=77)
        lat.set_path(path)
        lat.set_params([0.4, 1.0, 2.5], [0.0, 0.05, -0.1])
        if trial == "cone":
            lat.set_cone([0.5, 1.0, 2.0])
        lat.randomize()
        lat.equil(7)
        lat.equil(300, inst=[1])
        lats.append(lat)
    for i in range(3):
        assert np.array_equal(lats[0].get_spins(i), lats[1].get_spins(i)), i
    assert np.array_equal(lats[0].accepts(), lats[1].accepts())


@pytest.mark.parametrize("dim,periodic,masked", [(2, 1, False), (3, 0, True)])
def test_lds_ising_with_several_pairs_per_work_item(orc, bloch3d, dim, periodic, masked):
    # a 64-work-item LDS workgroup takes several pairs per colour (the loop
    # form, not the geometry kept per launch): still the streaming lattice
    rng = np.random.default_rng(47 + dim)
    edge = 10 if dim == 3 else 24
    bd, J, D = (orc.parse_bonds(bloch3d, 3, 1.0, 0.2) if dim == 3 else neel2d(0.2))
    locs = None
    if masked:
        locs = box_locs(dim, edge)
        locs = locs[rng.random(len(locs)) < 0.8]
    lats = []
    for path, block in (("auto", 64), ("auto", 0), ("stream", 0)):
        lat, f = make_pair(orc, dim, edge, periodic, bd, J, D, True, use_d=False, locs=locs, trial="sphere",
                           ninst=3, seed=78)
        lat.set_path(path)
        if block:
            lat.set_lds_block(block)
        lat.set_params([0.4, 1.0, 2.5], [0.0, 0.05, -0.1])
        lat.randomize()
        lat.equil(7)
        lat.equil(300, inst=[1])
        lats.append(lat)
    from nanomagmc_amd._lib import NMC_KIND_LDS
    assert lats[0].sweep_kinds().tolist() == [NMC_KIND_LDS] * 3
    for r in lats[1:]:
        for i in range(3):
            assert np.array_equal(lats[0].get_spins(i), r.get_spins(i)), i
        assert np.array_equal(lats[0].accepts(), r.accepts())


@pytest.mark.parametrize("kind", ["stream", "lds", "grains"])
def test_device_cone_adaptation_matches_host_steps(orc, bloch3d, kind):
    # nmc_sweep_adapt adapts the half-angles on the device after every chunk;
    # the same steps taken on the host (reset, sweep, read the accepts, new
    # angle, nmc_set_cone) must give the same lattices and angles bit for bit.
    # "stream": one 24^3 lattice (streamed, the captured-chunk path);
    # "lds": LDS-resident lattices; "grains": compacted and streamed grains.
    from nanomagmc_amd import Bond, Lattices
    rng = np.random.default_rng(13)
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    bonds = [Bond(bd[k], J[k], D[k]) for k in range(len(J))]
    if kind == "stream":
        insts, periodic = [{"n": (24, 24, 24), "edge": 24}], True
    elif kind == "lds":
        insts, periodic = [{"n": (8, 8, 8), "edge": 8}] * 3, True
    else:
        insts, periodic = [], False
        for n in (10, 7, 32):
            insts.append({"n": (n,) * 3, "edge": n, "mask": (rng.random(n ** 3) < 0.75).astype(np.uint8)})
    ni = len(insts)
    betas = list(np.linspace(0.3, 3.0, ni))
    lats = []
    for _ in range(2):
        lat = Lattices(3, False, periodic, bonds, insts, use_d=True, trial="cone", seed=57)
        lat.set_params(betas, 0.05)
        lat.set_cone([0.7] * ni)
        lat.randomize()
        lats.append(lat)
    sweeps, chunk = 125, 10   # two captured graphs of six chunks (streamed set), then 5 sweeps
    got = lats[0].equil_adapt(sweeps, chunk)
    host = lats[1]
    theta = np.full(ni, 0.7, dtype=np.float32)
    sites = [host.num_sites(i) for i in range(ni)]
    done = 0
    while done < sweeps:
        c = min(chunk, sweeps - done)
        host.reset_accepts()
        host.equil(c)
        acc = host.accepts()
        for i in range(ni):
            rate = float(acc[i]) / max(1.0, float(sites[i]) * c)
            f = min(2.0, max(0.5, rate / 0.5))
            theta[i] = np.float32(min(math.pi, max(0.02, float(theta[i]) * f)))
        host.set_cone(theta)
        done += c
    assert np.array_equal(got, theta)
    assert len(set(got.tolist())) > 1 or ni == 1
    for i in range(ni):
        assert np.array_equal(lats[0].get_spins(i), host.get_spins(i)), i
    assert np.all(lats[0].accepts() == 0)


@pytest.mark.parametrize("kind", ["periodic", "grains"])
def test_graphs_and_direct_launches_agree(orc, bloch3d, monkeypatch, kind):
    # a small streamed set replays captured graphs (plain sweeps, adapted
    # sweeps, the sample loop); NMC_NO_GRAPHS=1 launches the same kernels
    # directly: lattices and series must agree bit for bit.  "grains": open
    # grains swept compacted in LDS and from compacted copies, whose graphs hold
    # the two as branches on two streams.
    from nanomagmc_amd import Bond, Lattices
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    bonds = [Bond(bd[k], J[k], D[k]) for k in range(len(J))]
    rng = np.random.default_rng(61)
    grains = [{"n": (n,) * 3, "edge": n, "mask": (rng.random(n ** 3) < 0.75).astype(np.uint8)} for n in (9, 30)]
    runs = []
    for off in ("0", "1"):
        monkeypatch.setenv("NMC_NO_GRAPHS", off)
        if kind == "periodic":
            lat = Lattices(3, False, True, bonds, [{"n": (24, 24, 24), "edge": 24}] * 2, use_d=True, trial="cone",
                           seed=61)
        else:
            lat = Lattices(3, False, False, bonds, grains, use_d=True, trial="cone", seed=61)
            from nanomagmc_amd._lib import NMC_KIND_COMPACT_STREAM, NMC_KIND_LDS_COMPACT
            assert lat.sweep_kinds().tolist() == [NMC_KIND_LDS_COMPACT, NMC_KIND_COMPACT_STREAM]
        lat.set_params([0.5, 1.5], [0.0, 0.1])
        lat.randomize()
        lat.equil(150)
        th = lat.equil_adapt(130, 10)
        lat.series_reset(6, with_charge=True)
        lat.run_samples(6, 3)
        runs.append((lat, th, lat.get_series()))
    monkeypatch.delenv("NMC_NO_GRAPHS")
    assert np.array_equal(runs[0][1], runs[1][1])
    for i in range(2):
        assert np.array_equal(runs[0][0].get_spins(i), runs[1][0].get_spins(i)), i
    for a, b in zip(runs[0][2], runs[1][2]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dim,trial,table", [(3, "cone", "bloch"), (3, "sphere", "general"), (2, "sphere", "neel"),
                                              (2, "cone", "general")])
def test_compacted_and_streamed_grains_agree(bloch3d, orc, dim, trial, table):
    # one set holding grains small enough for the compacted LDS sweep and one
    # too large for it (streamed from its compacted copy): the lattice and the
    # accept counts must equal an all-streaming run over packed row spans,
    # and one streaming every grain from compacted copies, bit for bit, for
    # both DMI forms of the kernels
    from nanomagmc_amd import Bond, Lattices
    rng = np.random.default_rng(5 + dim)
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2) if dim == 3 else neel2d(0.2)
    D = np.asarray(D)
    if table == "general":
        D = D + 0.1 * np.sign(D.sum(axis=1, keepdims=True))
    bonds = [Bond(bd[k][:dim], J[k], D[k]) for k in range(len(J))]
    insts = []
    for n in ((10, 7, 32) if dim == 3 else (40, 23, 200)):
        m = (rng.random(n ** dim) < 0.75).astype(np.uint8)
        insts.append({"n": (n,) * dim, "edge": n, "mask": m})
    runs = []
    for path in ("auto", "stream", "cstream"):
        lat = Lattices(dim, False, False, bonds, insts, use_d=True, trial=trial, seed=91)
        lat.set_path(path)
        lat.set_params([0.7, 1.3, 2.0], [0.05, 0.0, -0.1])
        if trial == "cone":
            lat.set_cone([0.8, 1.1, 0.5])
        lat.randomize()
        lat.equil(5)
        lat.equil(260, inst=[0, 2])
        runs.append(lat)
    from nanomagmc_amd._lib import NMC_KIND_COMPACT_STREAM, NMC_KIND_LDS_COMPACT, NMC_KIND_SPANS
    assert runs[0].sweep_kinds().tolist() == [NMC_KIND_LDS_COMPACT, NMC_KIND_LDS_COMPACT, NMC_KIND_COMPACT_STREAM]
    assert runs[1].sweep_kinds().tolist() == [NMC_KIND_SPANS] * 3
    assert runs[2].sweep_kinds().tolist() == [NMC_KIND_COMPACT_STREAM] * 3
    for r in runs[1:]:
        for i in range(3):
            assert np.array_equal(runs[0].get_spins(i), r.get_spins(i)), i
        assert np.array_equal(runs[0].accepts(), r.accepts())


@pytest.mark.parametrize("ising", [True, False])
def test_graph_replays_match_eager_loop(orc, ising):
    # the captured sample loop, replayed several times with eager work in
    # between, gives exactly what the eager sweep + sample calls give
    d, J, D = nn2d() if ising else neel2d(0.3)
    series = []
    for mode in ("eager", "graph"):
        lat, f = make_pair(orc, 2, 32, 1, d, J, D, ising, use_d=not ising, ninst=2, seed=5)
        lat.set_params([0.5, 1.2], [0.0, 0.05])
        lat.randomize()
        lat.series_reset(6)
        out = []
        for rep in range(3):
            lat.equil(2)                     # eager work between replays
            if mode == "eager":
                for k in range(6):
                    lat.equil(3)
                    lat.sample(k)
            else:
                lat.run_samples(6, 3)
            out.append(lat.get_series()[0].copy())
        series.append(out)
    for a, b in zip(*series):
        assert np.array_equal(a, b)


# ------------------------------------------------------------ statistics ----
def test_ising_2d_energy_matches_kaufman(orc):
    # 2-D square Ising L=64 periodic, J=1, h=0 (the reference's CPU-runnable
    # case): <E> against Kaufman's exact finite-lattice energy, and against
    # Onsager's infinite-lattice u(T) within the stated finite-size term
    L = 64
    d, J, D = nn2d()
    Ts = [1.5, 2.0, 2.5, 3.5]
    lat, f = make_pair(orc, 2, L, 1, d, J, D, True, ninst=len(Ts))
    lat.set_params([1 / T for T in Ts], 0.0)
    lat.randomize()
    lat.equil(1000)
    nsamp = 1000
    lat.series_reset(nsamp, with_charge=False)
    for k in range(nsamp):
        lat.equil(10)
        lat.sample(k)
    obs, _, _ = lat.get_series()
    for i, T in enumerate(Ts):
        m, err, tau = exact.mean_err(obs[i, :, 9])
        want = exact.kaufman_energy(L, T)
        assert abs(m - want) < 4 * err + 1e-9, (T, m, want, err, tau)
        u = exact.onsager_energy_per_site(T)
        finite_size = abs(want / (L * L) - u)
        assert abs(m / (L * L) - u) < 4 * err / (L * L) + finite_size, (T, m / (L * L), u, finite_size)


@pytest.mark.parametrize("trial", ["sphere", "cone"])
def test_heisenberg_single_spin_and_dimer_langevin(orc, trial):
    # one spin in a field: <s_z> = L(beta h); open dimer: <E> = -J L(beta J)
    from nanomagmc_amd import Bond, Lattices
    bonds = [Bond((1, 0), 1.0), Bond((-1, 0), 1.0), Bond((0, 1), 1.0), Bond((0, -1), 1.0)]
    insts = [{"n": (1, 1), "edge": 2}] * 3 + [{"n": (2, 1), "edge": 2}] * 3
    lat = Lattices(2, False, False, bonds, insts, trial=trial, seed=11)
    betas = [0.5, 1.0, 2.0, 0.5, 1.0, 2.0]
    hs = [1.0, 1.0, 1.0, 0.0, 0.0, 0.0]
    lat.set_params(betas, hs)
    if trial == "cone":
        lat.set_cone(1.2)
    lat.randomize()
    lat.equil(200)
    nsamp = 20000
    lat.series_reset(nsamp, with_charge=False)
    for k in range(nsamp):
        lat.equil(2)
        lat.sample(k)
    obs, _, _ = lat.get_series()
    for i in range(3):
        m, err, _ = exact.mean_err(obs[i, :, 2])
        assert abs(m - exact.langevin(betas[i] * 1.0)) < 4 * err
    for i in range(3, 6):
        m, err, _ = exact.mean_err(obs[i, :, 9])
        assert abs(m - (-exact.langevin(betas[i]))) < 4 * err


@pytest.mark.parametrize("dim,form", [(2, "neel"), (2, "general"), (3, "bloch"), (3, "general")])
def test_dmi_dimer_energy(dim, form):
    # open dimer with exchange and DMI: <E_s> = -d ln Z / d beta of
    # E = -J s1.s2 - D.(s1 x s2).  "neel"/"bloch" tables put each bond's D on
    # one spin axis (the sweep kernels' two-term form), "general" tables do
    # not (the six-term form)
    from nanomagmc_amd import Bond, Lattices
    dirs = [(1, 0), (-1, 0), (0, 1), (0, -1)] if dim == 2 else \
        [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    d = 0.8
    if form == "neel":
        Dp = {0: (0.0, d, 0.0), 1: (-d, 0.0, 0.0)}           # z x r for +x, +y
    elif form == "bloch":
        Dp = {a: tuple(d * float(a == k) for k in range(3)) for a in range(3)}   # along the bond
    else:
        v = np.array([0.3, -0.4, 0.5])
        v *= d / np.linalg.norm(v)
        Dp = {a: tuple(np.roll(v, a)) for a in range(dim)}
    bonds = []
    for r in dirs:
        a = int(np.flatnonzero(r)[0])
        sgn = float(r[a])
        bonds.append(Bond(r, 1.0, tuple(sgn * x for x in Dp[a])))
    n = (2, 1) if dim == 2 else (2, 1, 1)
    betas = [0.5, 1.0, 2.0]
    lat = Lattices(dim, False, False, bonds, [{"n": n, "edge": 2}] * 3, use_d=True, seed=23)
    lat.set_params(betas, 0.0)
    lat.randomize()
    lat.equil(200)
    nsamp = 20000
    lat.series_reset(nsamp, with_charge=False)
    for k in range(nsamp):
        lat.equil(2)
        lat.sample(k)
    obs, _, _ = lat.get_series()
    for i, beta in enumerate(betas):
        m, err, _ = exact.mean_err(obs[i, :, 10])
        assert abs(m - exact.dmi_dimer_energy(beta, 1.0, d)) < 4 * err, (i, m, err)


def test_randomize_draws_uniform_directions(orc):
    from nanomagmc_amd import Bond, Lattices
    bonds = [Bond((1, 0, 0), 1.0), Bond((-1, 0, 0), 1.0)]
    lat = Lattices(3, False, True, bonds, [{"n": (32, 32, 32), "edge": 32}], seed=5)
    lat.randomize()
    s = lat.get_spins(0)
    nrm = np.linalg.norm(s, axis=1)
    assert np.all(np.abs(nrm - 1) < 2e-6)
    n = len(s)
    for a in range(3):
        assert abs(s[:, a].mean()) < 5 / math.sqrt(3 * n)
        assert abs((s[:, a] ** 2).mean() - 1 / 3) < 5 * math.sqrt(4 / 45 / n)
    # z uniform in (-1, 1): 10-bin chi^2
    hist, _ = np.histogram(s[:, 2], bins=10, range=(-1, 1))
    chi2 = ((hist - n / 10) ** 2 / (n / 10)).sum()
    assert chi2 < 40


def test_build_weibull_matches_oracle(orc):
    from nanomagmc_amd import build_weibull
    sizes = np.array([5.0, 7.5, 10.0])
    edge, wb, seed = 30, 8.0, 0x4E4D4304
    masks = build_weibull(3, edge, sizes, wb, seed)
    for g, a in enumerate(sizes):
        u = np.zeros(edge ** 3)
        for q in range((edge ** 3 + 3) // 4):
            w = orc.philox([q, 2 << 8, 0, 0], [seed & 0xFFFFFFFF, ((seed >> 32) + g) & 0xFFFFFFFF])
            u[4 * q:4 * q + 4] = (w / 4294967296.0)[: min(4, edge ** 3 - 4 * q)]
        locs = orc.build_sites(3, edge, "x", a, wb, u)
        want = np.zeros(edge ** 3, dtype=np.uint8)
        want[(locs[:, 0] * edge + locs[:, 1]) * edge + locs[:, 2]] = 1
        assert np.array_equal(masks[g], want)


def test_snapshot_series_and_average(orc):
    d, J, D = nn2d()
    lat, f = make_pair(orc, 2, 8, 1, d, J, D, False, ninst=2)
    lat.randomize()
    lat.equil(3, inst=[0])
    lat.snapshot([0], [1])
    assert np.array_equal(lat.get_spins(0), lat.get_spins(1))
    lat.zero_avg(0)
    acc = np.zeros((64, 3))
    for _ in range(3):
        lat.equil(1, inst=[0])
        lat.accumulate_avg(0)
        acc += lat.get_spins(0)
    assert np.allclose(lat.get_avg(0), acc, atol=1e-12)
    lat.series_reset(4)
    for k in range(4):
        lat.equil(1)
        lat.sample(k)
    obs, qr, qw = lat.get_series()
    o2, q2, w2 = lat.measure()
    assert np.array_equal(obs[:, 3, :], o2) and np.array_equal(qr[:, 3, :], q2)


@pytest.mark.parametrize("ising", [True, False])
@pytest.mark.parametrize("path", ["auto", "stream"])
def test_long_sweep_calls_replay_graphs_exactly(orc, ising, path):
    # a small set swept 200 times in one call (captured graphs of 64 sweeps
    # plus the remainder) equals 200 single-sweep calls bit for bit
    d, J, D = nn2d()
    spins, accs = [], []
    for one_call in (True, False):
        lat, f = make_pair(orc, 2, 32, 1, d, J, D, ising, ninst=2, seed=29)
        lat.set_path(path)
        lat.set_params([0.5, 1.2], 0.05)
        lat.randomize()
        if one_call:
            lat.equil(200)
        else:
            for _ in range(200):
                lat.equil(1)
        spins.append([lat.get_spins(i) for i in range(2)])
        accs.append(lat.accepts())
    for a, b in zip(*spins):
        assert np.array_equal(a, b)
    assert np.array_equal(accs[0], accs[1])


def test_batched_average_equals_single_calls(orc):
    d, J, D = nn2d()
    sums = []
    for batched in (True, False):
        lat, f = make_pair(orc, 2, 8, 1, d, J, D, False, ninst=3, seed=17)
        lat.randomize()
        for _ in range(4):
            lat.equil(1)
            if batched:
                lat.accumulate_avgs([2, 0])
            else:
                lat.accumulate_avg(0)
                lat.accumulate_avg(2)
        sums.append([lat.get_avg(i) for i in (0, 2)])
    for a, b in zip(*sums):
        assert np.array_equal(a, b)


def test_invalid_arguments_raise(orc):
    from nanomagmc_amd import Bond, Lattices
    from nanomagmc_amd._lib import NmcError
    bonds = [Bond((1, 0), 1.0), Bond((-1, 0), 1.0)]
    lat = Lattices(2, True, True, bonds, [{"n": (4, 4), "edge": 4}])
    with pytest.raises(NmcError):
        lat.set_params(-1.0, 0.0)          # T <= 0 (reference exit 104)
    lat2 = Lattices(2, True, True, bonds, [{"n": (4, 4), "edge": 4}] * 2)
    with pytest.raises(NmcError):
        lat2.equil(1, inst=[1, 0, 1])      # one instance twice in a sweep list
    with pytest.raises(NmcError):
        lat2.accumulate_avgs([0, 0])
    with pytest.raises(NmcError):
        Lattices(2, True, True, bonds, [{"n": (5, 5), "edge": 5}])   # odd periodic edge
    with pytest.raises(NmcError):
        Lattices(2, True, False, [Bond((2, 0), 1.0)], [{"n": (4, 4), "edge": 4}])
    with pytest.raises(NmcError):
        lat2.equil_adapt(10, 10)           # cone adaptation needs Heisenberg spins
    heis = Lattices(2, False, True, bonds, [{"n": (4, 4), "edge": 4}], trial="cone")
    with pytest.raises(NmcError):
        heis.equil_adapt(10, 0)            # chunk must be > 0
    with pytest.raises(NmcError):
        heis.set_cone([0.0])               # half-angle outside (0, pi]
    assert heis.equil_adapt(0, 10).shape == (1,)


@pytest.mark.parametrize("table,trial", [("neel", "sphere"), ("general", "cone")])
def test_fused_sample_loop_2d_equals_the_calls(table, trial):
    # the fused sample loop on 2-D open grains (one compacted in LDS, one
    # streamed from its compacted copy), axial and general DMI, sphere and
    # cone trials: series and lattices equal the step-by-step loop's
    from nanomagmc_amd import Bond, Lattices
    from nanomagmc_amd._lib import NMC_KIND_COMPACT_STREAM, NMC_KIND_LDS_COMPACT
    d, J, D = neel2d(0.3)
    D = np.asarray(D)
    if table == "general":
        D = D + 0.1 * np.sign(D.sum(axis=1, keepdims=True))
    bonds = [Bond(d[k][:2], J[k], D[k]) for k in range(4)]
    rng = np.random.default_rng(17)
    insts = [{"n": (n, n), "edge": n, "mask": (rng.random(n * n) < 0.8).astype(np.uint8)} for n in (41, 180)]
    lats = []
    for way in ("loop", "calls"):
        lat = Lattices(2, False, False, bonds, insts, use_d=True, trial=trial, seed=23)
        lat.set_params([0.8, 2.5], [0.06, 0.0])
        if trial == "cone":
            lat.set_cone([0.9, 0.6])
        lat.randomize()
        lat.equil(4)
        lat.series_reset(5, with_charge=False)
        if way == "loop":
            lat.run_samples(5, 2)
        else:
            for slot in range(5):
                lat.equil(2)
                lat.sample(slot)
        lats.append(lat)
    a, b = lats
    assert a.sweep_kinds().tolist() == [NMC_KIND_LDS_COMPACT, NMC_KIND_COMPACT_STREAM]
    for x, y in zip(a.get_series(), b.get_series()):
        assert np.array_equal(x, y)
    for i in range(2):
        assert np.array_equal(a.get_spins(i), b.get_spins(i)), i


@pytest.mark.parametrize("graphs,charge", [("0", True), ("1", True), ("0", False), ("1", False)])
def test_sample_loop_with_averages_equals_the_calls(orc, bloch3d, monkeypatch, graphs, charge):
    # nmc_run_samples_avg (the drop-in driver's whole sample loop, one graph)
    # against nmc_sweep + nmc_sample + nmc_accumulate_avgs per slot, on open
    # grains (compacted in LDS and streamed from compacted copies).  Without
    # charges the loop runs fused: the LDS-resident grains sweep and measure
    # in one launch, the streamed one is measured from its compacted copy;
    # series, averages, lattices and accept counts must still be the calls'
    # bit for bit.
    from nanomagmc_amd import Bond, Lattices
    monkeypatch.setenv("NMC_NO_GRAPHS", "1" if graphs == "0" else "0")
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    bonds = [Bond(bd[k], J[k], D[k]) for k in range(len(J))]
    rng = np.random.default_rng(71)
    insts = [{"n": (n,) * 3, "edge": n, "mask": (rng.random(n ** 3) < 0.75).astype(np.uint8)} for n in (9, 30, 11)]
    lats = []
    for way in ("loop", "calls"):
        lat = Lattices(3, False, False, bonds, insts, use_d=True, seed=71)
        lat.set_params([0.5, 1.5, 3.0], [0.0, 0.1, 0.05])
        lat.randomize()
        lat.equil(20)
        lat.series_reset(7, with_charge=charge)
        avg = [0, 2] if charge else [0, 1, 2]
        for i in avg:
            lat.zero_avg(i)
        lat.reset_accepts()
        if way == "loop":
            lat.run_samples_avg(7, 3, avg)
        else:
            for slot in range(7):
                lat.equil(3)
                lat.sample(slot)
                lat.accumulate_avgs(avg)
        lats.append(lat)
    a, b = lats
    from nanomagmc_amd._lib import NMC_KIND_COMPACT_STREAM, NMC_KIND_LDS_COMPACT
    assert a.sweep_kinds().tolist() == [NMC_KIND_LDS_COMPACT, NMC_KIND_COMPACT_STREAM, NMC_KIND_LDS_COMPACT]
    for x, y in zip(a.get_series(), b.get_series()):
        assert np.array_equal(x, y)
    assert np.abs(a.get_series()[0][:, :, 10]).min() > 0
    for i in range(3):
        assert np.array_equal(a.get_spins(i), b.get_spins(i)), i
    for i in avg:
        assert np.array_equal(a.get_avg(i), b.get_avg(i)), i
        assert np.abs(a.get_avg(i)).sum() > 0
    assert np.array_equal(a.accepts(), b.accepts())
    # the loop continues the same random streams: one more sweep of each
    # lattice stays equal
    a.equil(2)
    b.equil(2)
    for i in range(3):
        assert np.array_equal(a.get_spins(i), b.get_spins(i)), i
