"""The grain batch (512 Weibull spheres x 64 (T, h), open boundaries, Bloch
DMI) and the full-tile launches of the 3-D Heisenberg batch, against the
oracle.  Needs an MI355X.

* Sampled observables of single Weibull grains swept by the kernels the grain
  batch uses -- compacted LDS-resident (AUTO) and streamed from a compacted
  copy in HBM (CSTREAM, the kernel of the batch's large grains) -- and by the
  packed row-span streaming kernel (STREAM), and of a periodic masked
  lattice (the masked LDS-resident kernel), against the oracle's FP64
  random-site Metropolis
  chain on the same site list (the reference's state::equil with dE_func on
  a shape::weibull lattice, Old/lib/state.cpp:124-203, shape.cpp:40-58):
  <E_s>, <|M|>, C and chi at 3 sigma, errors from the spread of independent
  replicas (so tau_int is accounted for).
* The whole 32768-instance grain batch swept on the device: measurements of
  16 instances (the largest, streamed grains among them) against the
  oracle's energy, magnetisations and charges of the spins read back, and
  unit spins everywhere (a non-zero absent site would enter the oracle-
  compared energies of its present neighbours).
* Full-tile streaming launches (256 work-items, 1024-element tiles: the
  timed launch shape of the 64^3 batch) against single-instance quarter-
  tile launches of the same chains, bit for bit (a site's random numbers do
  not depend on the launch shape), and the 64^3 measurement against the
  oracle; full-tile Ising passes against the oracle's checkerboard sweep,
  bit for bit.
"""
import math

import numpy as np
import pytest

from tests.fixtures import box_locs, neel2d, nn2d, random_unit
from tests.gpu_util import estimators, make_pair, oracle_chain, replica_stats

pytestmark = pytest.mark.gpu


def grain(dim, edge, size, seed):
    """One Weibull grain (WEIBULLFACT 8) built on the device: (instance
    description, site coordinates in the edge box in site order)."""
    from nanomagmc_amd import build_weibull, crop_to_box
    m = build_weibull(dim, edge, [size], 8.0, seed=seed)[0]
    n, lo, sub = crop_to_box(m, dim, edge)
    idx = np.argwhere(m.reshape((edge,) * dim))      # box-scan order, last coordinate fastest
    locs = np.zeros((len(idx), 3), dtype=np.int32)
    locs[:, :dim] = idx
    return {"n": tuple(n), "origin": tuple(lo), "edge": edge, "mask": sub}, locs


GRAINS = {
    # dim, edge, radius scale, table, D, h, T
    "bloch3d": (3, 11, 4.0, "bloch", 0.2, 0.05, 1.0),
    "neel2d": (2, 15, 6.0, "neel", 0.3, 0.06, 0.5),
}

_chains = {}


def _oracle_stats(orc, key, f, H, beta, N, nchains=16, ns=2000):
    if key not in _chains:
        rng = np.random.default_rng(0x4E4D4320 + len(key))
        ref = []
        for _ in range(nchains):
            E, M, _ = oracle_chain(orc, f, H, beta, ns, 5, rng)
            ref.append(estimators(E, M, beta, N))
        _chains[key] = ref
    return _chains[key]


def _table(orc, bloch3d, kind, D):
    if kind == "bloch":
        return orc.parse_bonds(bloch3d, 3, 1.0, D)
    return neel2d(D)


@pytest.mark.parametrize("which,path,trial", [("bloch3d", "auto", "sphere"), ("bloch3d", "stream", "sphere"),
                                              ("bloch3d", "cstream", "sphere"), ("bloch3d", "auto", "cone"),
                                              ("neel2d", "auto", "sphere"), ("neel2d", "stream", "sphere"),
                                              ("neel2d", "cstream", "sphere")])
def test_weibull_grain_observables_match_oracle_chain(orc, bloch3d, which, path, trial):
    from nanomagmc_amd import Bond, Lattices
    dim, edge, size, kind, D, h, T = GRAINS[which]
    inst, locs = grain(dim, edge, size, seed=0x4E4D4330 + dim)
    N = len(locs)
    assert 60 < N < 400
    bd, Jv, Dv = _table(orc, bloch3d, kind, D)
    bonds = [Bond(bd[k][:dim], Jv[k], Dv[k]) for k in range(len(Jv))]
    beta = 1.0 / T
    R, ns = 32, 4000
    lat = Lattices(dim, False, False, bonds, [dict(inst) for _ in range(R)], use_d=True, trial=trial,
                   seed=0x4E4D4331 + dim)
    lat.set_path(path)
    lat.set_params([beta] * R, h)
    lat.randomize()
    if trial == "cone":
        lat.set_cone([1.0] * R)
        lat.equil_adapt(400, 10)
    lat.equil(500)
    lat.series_reset(ns, with_charge=False)
    lat.run_samples(ns, 5)
    obs, _, _ = lat.get_series()
    assert lat.num_sites(0) == N
    dev = [estimators(obs[i, :, 10], obs[i, :, 0:3], beta, N) for i in range(R)]
    f = orc.Field(locs, dim, edge, 0, bd, Jv, Dv)
    ref = _oracle_stats(orc, which, f, [0.0, 0.0, h], beta, N)
    for k in ("E", "absM", "C", "chi"):
        m, e = replica_stats([p[k] for p in dev])
        mr, er = replica_stats([p[k] for p in ref])
        sig = math.hypot(e, er)
        print("%s %s %s %-5s device %.5f +- %.5f   oracle %.5f +- %.5f   (%.2f sigma)"
              % (which, path, trial, k, m, e, mr, er, abs(m - mr) / sig))
        assert abs(m - mr) < 3 * sig, (k, m, e, mr, er)


def test_periodic_masked_lattice_observables_match_oracle_chain(orc):
    # a periodic lattice with a random site mask fits in LDS with its mask
    # words: the masked LDS-resident kernel (AUTO) against the oracle chain
    from nanomagmc_amd import Bond, Lattices
    L, D, h, T = 8, 0.3, 0.06, 0.6
    rng = np.random.default_rng(0x4E4D4332)
    locs = box_locs(2, L)
    locs = locs[rng.random(len(locs)) < 0.8]
    d, Jv, Dv = neel2d(D)
    beta = 1.0 / T
    R, ns = 32, 4000
    lat, f = make_pair(orc, 2, L, 1, d, Jv, Dv, False, use_d=True, locs=locs, ninst=R, seed=0x4E4D4333)
    lat.set_params([beta] * R, h)
    lat.randomize()
    lat.equil(500)
    lat.series_reset(ns, with_charge=False)
    lat.run_samples(ns, 5)
    obs, _, _ = lat.get_series()
    N = f.n
    dev = [estimators(obs[i, :, 10], obs[i, :, 0:3], beta, N) for i in range(R)]
    ref = _oracle_stats(orc, "periodic_masked", f, [0.0, 0.0, h], beta, N)
    for k in ("E", "absM", "C", "chi"):
        m, e = replica_stats([p[k] for p in dev])
        mr, er = replica_stats([p[k] for p in ref])
        sig = math.hypot(e, er)
        print("periodic masked %-5s device %.5f +- %.5f   oracle %.5f +- %.5f   (%.2f sigma)"
              % (k, m, e, mr, er, abs(m - mr) / sig))
        assert abs(m - mr) < 3 * sig, (k, m, e, mr, er)


def test_full_grain_batch_against_oracle(orc, bloch3d):
    from nanomagmc_amd.workloads import grain_masks, grains
    sizes, masks = grain_masks()
    lat, sites, _, _ = grains("sphere")
    assert lat.ninst == 32768 and len(sites) == 32768
    lat.equil(3)
    obs, qr, qw = lat.measure()
    assert np.all(np.isfinite(obs))
    edge = 45
    bd, Jv, Dv = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    nk = np.array([int(m.sum()) for m in masks])
    order = np.argsort(nk)
    # the 8 largest grains (streamed over their quads) and 8 spread over the rest
    picks = list(order[-8:]) + list(order[np.linspace(0, len(order) - 9, 8).astype(int)])
    temps = np.linspace(2.0, 0.1, 16)
    hs = [0.0, 0.02, 0.05, 0.1]
    for j, g in enumerate(picks):
        i = 64 * int(g) + (7 * j + 3) % 64
        h = hs[(i % 64) // 16]
        s = lat.get_spins(i)
        assert s.shape[0] == nk[g] == sites[i]
        assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=2e-6), i
        idx = np.argwhere(masks[g].reshape(edge, edge, edge))
        f = orc.Field(idx.astype(np.int32), 3, edge, 0, bd, Jv, Dv)
        H = [0.0, 0.0, h]
        want = np.concatenate([f.mag(s), f.submag(s), f.sub4mag(s),
                               [f.energy(s, H, use_d=True, convention=0), f.energy(s, H, use_d=True, convention=1)]])
        assert np.allclose(obs[i], want, rtol=1e-12, atol=1e-9), (g, i, nk[g])
        assert np.allclose(qr[i], f.tcharge(s, wrap=False), rtol=1e-10, atol=1e-10), (g, i)
        print("grain %d (%d sites) instance %d T=%.3f h=%.2f: E_s %.6f" % (g, nk[g], i, temps[i % 16], h, obs[i, 10]))


def test_full_tile_heisenberg_launches_equal_quarter_tiles(orc):
    # 8 instances of 64^3 in one set fill 1024 full tiles (256 work-items),
    # the launch shape of the timed batch; each instance swept in a set of
    # its own takes quarter tiles (64 work-items).  Same random numbers per
    # site, so the lattices must agree bit for bit.
    from nanomagmc_amd.workloads import heis3d
    a, _, _, _ = heis3d("cone", ninst=8)
    b, _, _, _ = heis3d("cone", ninst=8)
    cones = np.linspace(0.4, 2.5, 8).astype(np.float32)
    for lat in (a, b):
        lat.set_cone(cones)
    a.equil(3)
    for i in range(8):
        b.equil(3, inst=[i])
    for i in range(8):
        assert np.array_equal(a.get_spins(i), b.get_spins(i)), i
    assert np.array_equal(a.accepts(), b.accepts())
    # and the measurement of one of them against the oracle
    s = a.get_spins(5)
    bd = np.array([[0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0]], dtype=np.int32)
    f = orc.Field(box_locs(3, 64), 3, 64, 1, bd, np.ones(6), np.zeros((6, 3)))
    obs, qr, qw = a.measure([5])
    h = 0.0
    want = np.concatenate([f.mag(s), f.submag(s), f.sub4mag(s),
                           [f.energy(s, [0, 0, h], convention=0), f.energy(s, [0, 0, h], convention=1)]])
    assert np.allclose(obs[0], want, rtol=1e-12, atol=1e-9)
    assert np.allclose(qr[0], f.tcharge(s, wrap=False), rtol=1e-10, atol=1e-10)
    assert np.allclose(qw[0], f.tcharge(s, wrap=True), rtol=1e-10, atol=1e-10)


def test_full_tile_ising_passes_match_oracle_sweeps(orc):
    # 2-D Ising L = 1024: 512 Ki elements per colour = 1024 full tiles of 512
    # (256 work-items); two sweeps against the oracle's checkerboard sweep
    # with the same Philox stream, bit for bit
    L, T, h, seed = 1024, 2.3, 0.05, 0x4E4D4334
    d, J, D = nn2d()
    lat, f = make_pair(orc, 2, L, 1, d, J, D, True, seed=seed)
    rng = np.random.default_rng(9)
    s = np.zeros((L * L, 3))
    s[:, 0] = rng.choice([-1.0, 1.0], L * L)
    lat.set_spins(0, s)
    lat.set_params(1 / T, h)
    lat.equil(2)
    e, c = f.colour_index((L, L))
    f.checkerboard_ising(s, [h, 0, 0], 1 / T, e, c, seed=seed, inst=0, sweep0=0, nsweeps=2)
    assert np.array_equal(lat.get_spins(0), s)
    obs, _, _ = lat.measure()
    assert obs[0, 9] == f.energy(s, [h, 0, 0])


def test_lds_capacity_boundary_and_degenerate_grains(orc, bloch3d):
    # open 3-D grains in a 32^3 box at the compacted kernels' limits:
    # 13,567 present sites (the most one workgroup's LDS holds on MI355X:
    # 12 B a site plus the zero slot, with 1 KB to spare) and one more
    # (streamed from a compacted copy instead), a single site, 40 sites of
    # one colour only, and no site at all.  AUTO, STREAM (packed row spans)
    # and CSTREAM give the same lattices and accept counts bit for bit, the
    # fused sample loop the step-by-step loop's series; the lone site's
    # energy is its Zeeman term alone, the empty grain's observables zero.
    from nanomagmc_amd import Bond, Lattices
    from nanomagmc_amd._lib import NMC_KIND_COMPACT_STREAM, NMC_KIND_LDS_COMPACT, NMC_KIND_SPANS
    n = 32
    rng = np.random.default_rng(41)
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    bonds = [Bond(bd[k][:3], J[k], D[k]) for k in range(len(J))]

    def mask_of(sites):
        m = np.zeros(n ** 3, np.uint8)
        m[sites] = 1
        return m

    idx = np.arange(n ** 3)                          # box-scan order, last coordinate fastest
    even = np.flatnonzero((idx // (n * n) + (idx // n) % n + idx % n) % 2 == 0)
    perm = rng.permutation(n ** 3)
    masks = [mask_of(perm[:13567]), mask_of(perm[:13568]), mask_of(perm[:1]),
             mask_of(rng.choice(even, 40, replace=False)), np.zeros(n ** 3, np.uint8)]
    insts = [{"n": (n, n, n), "edge": n, "mask": m} for m in masks]
    beta, h = [1.1, 1.0, 2.0, 0.7, 1.5], [0.05, 0.0, 0.1, -0.05, 0.02]
    runs = []
    for path in ("auto", "stream", "cstream"):
        lat = Lattices(3, False, False, bonds, insts, use_d=True, trial="sphere", seed=57)
        lat.set_path(path)
        lat.set_params(beta, h)
        lat.randomize()
        lat.equil(3)
        lat.equil(30)
        runs.append(lat)
    a = runs[0]
    assert [a.num_sites(i) for i in range(5)] == [13567, 13568, 1, 40, 0]
    assert a.sweep_kinds().tolist() == [NMC_KIND_LDS_COMPACT, NMC_KIND_COMPACT_STREAM] + [NMC_KIND_LDS_COMPACT] * 3
    assert runs[1].sweep_kinds().tolist() == [NMC_KIND_SPANS] * 5
    assert runs[2].sweep_kinds().tolist() == [NMC_KIND_COMPACT_STREAM] * 5
    for r in runs[1:]:
        for i in range(4):
            assert np.array_equal(a.get_spins(i), r.get_spins(i)), i
        assert np.array_equal(a.accepts(), r.accepts())
    acc = a.accepts()
    assert acc[4] == 0 and acc[2] > 0
    for i in range(4):
        s = a.get_spins(i)
        assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-6)
    obs, _, _ = a.measure()
    s2 = a.get_spins(2)[0]
    assert obs[2, 10] == pytest.approx(-h[2] * s2[2], abs=1e-12)
    assert obs[2, 9] == pytest.approx(-h[2] * s2[2], abs=1e-12)
    assert np.all(obs[4] == 0.0)
    # the fused sample loop (every grain compacted, no charges) against the
    # step-by-step loop, from the same state
    ways = []
    for way in ("loop", "calls"):
        lat = Lattices(3, False, False, bonds, insts, use_d=True, trial="sphere", seed=57)
        lat.set_params(beta, h)
        lat.randomize()
        lat.equil(3)
        lat.series_reset(4, with_charge=False)
        if way == "loop":
            lat.run_samples(4, 3)
        else:
            for slot in range(4):
                lat.equil(3)
                lat.sample(slot)
        ways.append(lat)
    for x, y in zip(ways[0].get_series(), ways[1].get_series()):
        assert np.array_equal(x, y)
    assert np.all(ways[0].get_series()[0][4] == 0.0)
    for i in range(4):
        assert np.array_equal(ways[0].get_spins(i), ways[1].get_spins(i)), i


def test_size_classes_long_and_short_sweep_calls_agree(orc, bloch3d):
    # open 3-D grains whose larger colour's pair count m falls in every size
    # class of the compacted kernel: (0, 128], (128, 256], (256, 512],
    # (512, 1024] and above.  A call of 40 sweeps launches the upper two
    # classes apart (two pairs per work-item in (512, 1024]); calls of 5
    # sweeps launch them together with full-cover workgroups, and a forced
    # workgroup size launches one class.  The lattices and accept counts
    # agree bit for bit: a site's random numbers do not depend on the launch.
    from nanomagmc_amd import Bond, Lattices
    from nanomagmc_amd._lib import NMC_KIND_LDS_COMPACT
    n = 32
    rng = np.random.default_rng(43)
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    bonds = [Bond(bd[k][:3], J[k], D[k]) for k in range(len(J))]
    idx = np.arange(n ** 3)
    colour = (idx // (n * n) + (idx // n) % n + idx % n) % 2
    masks, ms = [], []
    for count in (200, 400, 600, 1500, 2400, 4000, 9000):
        m = np.zeros(n ** 3, np.uint8)
        m[rng.permutation(n ** 3)[:count]] = 1
        # pair k of a colour holds its colour elements 2k and 2k + 1 (box-scan order)
        mc = [int(np.count_nonzero(m[colour == c].reshape(-1, 2).any(axis=1))) for c in (0, 1)]
        masks.append(m)
        ms.append(max(mc))
    cls = [0 if x <= 128 else 1 if x <= 256 else 2 if x <= 512 else 3 if x <= 1024 else 4 for x in ms]
    assert sorted(set(cls)) == [0, 1, 2, 3, 4], ms
    insts, beta, h = [], [], []
    for k, m in enumerate(masks):
        for j in range(3):
            insts.append({"n": (n, n, n), "edge": n, "mask": m})
            beta.append(0.6 + 0.4 * j)
            h.append(0.05 * (k % 3))
    runs = []
    for way in ("long", "short", "forced"):
        lat = Lattices(3, False, False, bonds, insts, use_d=True, trial="sphere", seed=61)
        if way == "forced":
            lat.set_lds_block(1024)
        lat.set_params(beta, h)
        lat.randomize()
        if way == "short":
            for _ in range(8):
                lat.equil(5)
        else:
            lat.equil(40)
        runs.append(lat)
    a = runs[0]
    assert a.sweep_kinds().tolist() == [NMC_KIND_LDS_COMPACT] * len(insts)
    for r in runs[1:]:
        for i in range(len(insts)):
            assert np.array_equal(a.get_spins(i), r.get_spins(i)), i
        assert np.array_equal(a.accepts(), r.accepts())
    assert np.all(a.accepts() > 0)
