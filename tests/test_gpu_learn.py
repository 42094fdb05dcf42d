"""Learned block proposals on the device: gradients of the network against
the FP64 oracle, the Metropolis-Hastings quantities against oracle energies,
exact stationarity on a solvable case, training, and the effect on
autocorrelation times.  Needs an MI355X."""
import math

import numpy as np
import pytest

from tests import exact
from tests.fixtures import box_locs, neel2d, random_unit

pytestmark = pytest.mark.gpu


def lattice(L, D, J=1.0, ninst=1, seed=9):
    from nanomagmc_amd import Bond, Lattices
    d, Jv, Dv = neel2d(D)
    bonds = [Bond(d[k], J * Jv[k], Dv[k]) for k in range(4)]
    return Lattices(2, False, True, bonds, [{"n": (L, L), "edge": L}] * ninst, use_j=J != 0, use_d=D != 0,
                    seed=seed)


def last_bias_offset():
    """Offset of the output layer's bias (mu_x, mu_y, mu_z, kappa pre-activation)."""
    from oracle import netref
    return sum(9 * c * 32 + 32 for c in netref.CIN[:3]) + 9 * netref.CIN[3] * 32


def patches_of(spins, L, blocks):
    """The device's network input for blocks at (bx, by) (the oracle's
    patch(): halo spins and the block's interpolated ring) and the block
    targets."""
    from oracle import netref
    s = spins.reshape(L, L, 3)
    P = np.stack([netref.patch(spins, L, bx, by) for bx, by in blocks])
    T = np.stack([s[np.ix_(np.arange(bx, bx + 16) % L, np.arange(by, by + 16) % L)].reshape(256, 3)
                  for bx, by in blocks])
    return P, T


def test_network_gradients_match_oracle():
    from nanomagmc_amd.learn import Net
    from oracle import netref
    rng = np.random.default_rng(3)
    net = Net(seed=1)
    p = netref.init_params(rng).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    n = 6
    P = np.zeros((n, 24, 24, 4))
    s = random_unit(n * 576, rng).reshape(n, 24, 24, 3)
    P[..., :3] = s
    P[..., 3] = 1.0
    P[:, 4:20, 4:20, :] = 0.0
    T = random_unit(n * 256, rng).reshape(n, 256, 3)
    loss, g = net.loss_grad(P, T)
    ref_loss, ref_g = netref.loss_grad(p.astype(np.float64), P, T, use_bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    off = 0
    for l, c in enumerate(netref.CIN):
        K = 9 * c
        for part, size in (("W", K * 32), ("b", 32)):
            a, b = g[off:off + size], ref_g[off:off + size]
            off += size
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
            assert err < 3e-2, (l, part, err)


@pytest.mark.parametrize("T,kappa0", [(2.0, None), (100.0, 0.05)])
def test_block_move_quantities_match_oracle(orc, T, kappa0):
    # kappa0 = None: a random network at T = 2 (almost every block rejected);
    # kappa0 = 0.05: a near-uniform constant proposal at T = 100, where
    # accepted and rejected blocks mix
    from nanomagmc_amd.learn import Net, block_moves
    from oracle import netref
    L, D, h = 64, 0.3, 0.06
    lat = lattice(L, D)
    lat.set_params(1 / T, h)
    rng = np.random.default_rng(5)
    before = random_unit(L * L, rng).astype(np.float32).astype(np.float64)
    lat.set_spins(0, before)
    net = Net(seed=2)
    p = netref.init_params(rng).astype(np.float32)
    if kappa0 is not None:
        p[:] = 0.0
        p[last_bias_offset():last_bias_offset() + 4] = [0.3, -0.2, 0.9, math.log(math.expm1(kappa0 - 0.05 + 1e-3))]
    net.set_params(p)
    off = (7, 19)
    acc, stats, prop = block_moves(lat, net, [0], *off, diagnostics=True)
    after = lat.get_spins(0)
    blocks = [(32 * i + off[0], 32 * j + off[1]) for i in range(2) for j in range(2)]
    d, Jv, Dv = neel2d(D)
    f = orc.Field(box_locs(2, L), 2, L, 1, d, Jv, Dv)
    H = [0, 0, h]
    E0 = f.energy(before, H, use_d=True, convention=1)
    P, Tg = patches_of(before, L, blocks)
    _, O = netref.forward(p.astype(np.float64), P, use_bf16=True)
    _, kappa, _ = netref.head(O)
    mu = O[..., :3] / np.linalg.norm(O[..., :3], axis=-1, keepdims=True)   # proposals use the exact unit mean
    changed = np.zeros(L * L, dtype=bool)
    naccepted = 0
    for b, (bx, by) in enumerate(blocks):
        xs = np.arange(bx, bx + 16) % L
        ys = np.arange(by, by + 16) % L
        idx = (xs[:, None] * L + ys[None, :]).reshape(-1)
        trial = before.copy()
        trial[idx] = prop[b].reshape(256, 3)
        assert np.allclose(np.linalg.norm(trial[idx], axis=1), 1.0, atol=1e-6)
        dE = f.energy(trial, H, use_d=True, convention=1) - E0
        assert abs(stats[b, 0] - dE) <= 2e-3 * (1 + abs(dE)), (b, stats[b, 0], dE)
        lqn = netref.logq(mu[b], kappa[b], prop[b].astype(np.float64)).sum()
        lqo = netref.logq(mu[b], kappa[b], Tg[b].reshape(16, 16, 3)).sum()
        assert abs(stats[b, 1] - lqn) <= 0.05 * (1 + abs(lqn)), (b, stats[b, 1], lqn)
        assert abs(stats[b, 2] - lqo) <= 0.05 * (1 + abs(lqo)), (b, stats[b, 2], lqo)
        # the decision: thread 0's Philox word 2 at counter (b * 256, block purpose, first block move)
        w = orc.philox([b * 256, 3 << 8, 1, 0], [9, 0])
        u = np.float32(((int(w[2]) >> 8) + 0.5) / 16777216.0)
        loga = np.float32(-np.float32(1 / T) * stats[b, 0] + stats[b, 2] - stats[b, 1])
        accept = bool(loga >= 0 or np.log(u) < loga)
        now = after[idx]
        if accept:
            assert np.array_equal(now, prop[b].reshape(256, 3).astype(np.float64))
            naccepted += 1
        else:
            assert np.array_equal(now, before[idx])
        changed[idx] = True
    assert acc[0] == naccepted
    assert np.array_equal(after[~changed], before[~changed])   # nothing outside the blocks moves


def test_block_moves_alone_sample_the_exact_distribution():
    # independent spins in a field (J = D = 0): <s_z> = L(beta h).  The
    # network is set to a constant vMF(z, kappa0) with kappa0 != beta h, so
    # only the Hastings correction keeps the chain exact.  Plain sweeps
    # equilibrate first (an independence proposal on 256 spins accepts
    # nothing from a uniform start); after that only block moves act.
    from nanomagmc_amd.learn import Net, block_moves
    from oracle import netref
    L, h, beta = 64, 1.0, 1.0
    lat = lattice(L, 0.0, J=0.0)
    lat.set_params(beta, h)
    lat.randomize()
    lat.equil(200)
    net = Net(seed=3)
    p = np.zeros(netref.n_params(), dtype=np.float32)
    kappa0 = 1.1
    z = math.log(math.expm1(kappa0 - 0.05))
    off = last_bias_offset()
    p[off:off + 4] = [0.0, 0.0, 1.0, z]           # last-layer bias: mu = z-axis, kappa = kappa0
    net.set_params(p)
    rng = np.random.default_rng(11)
    n = 1500
    lat.series_reset(n, with_charge=False)
    nacc = 0
    for k in range(n):
        for _ in range(2):
            ox, oy = rng.integers(0, 32, 2)
            nacc += int(block_moves(lat, net, [0], int(ox), int(oy))[0])
        lat.sample(k)
    obs, _, _ = lat.get_series()
    sz = obs[0, 100:, 2] / (L * L)
    m, err, _ = exact.mean_err(sz)
    rate = nacc / (n * 2 * 4)
    print("block acceptance %.3f, <s_z> = %.5f +- %.5f, exact %.5f" % (rate, m, err, exact.langevin(beta * h)))
    assert 0.05 < rate < 0.95
    assert abs(m - exact.langevin(beta * h)) < 4 * err, (m, exact.langevin(beta * h), err)
    assert abs(exact.langevin(kappa0) - exact.langevin(beta * h)) > 20 * err   # the test can see a bias


def test_training_lowers_the_loss_and_keeps_the_physics():
    # T = 0.6: short correlation length, so two 64 x 64 instances mix fast
    # enough that their mean energies must agree within the error bars
    from nanomagmc_amd.learn import Net, block_moves
    L, D, h, T = 64, 0.3, 0.06, 0.6
    lat = lattice(L, D, ninst=2, seed=21)
    lat.set_params(1 / T, h)
    lat.randomize()
    lat.equil(500)
    net = Net(seed=4)
    for k in range(40):
        lat.equil(10)
        net.collect(lat, 0, 64, seed=k)
    assert net.num_samples() == 40 * 64
    first = net.train(1, batch=128, lr=0.0, seed=1)
    last = net.train(300, batch=128, lr=2e-3, seed=2)
    assert last < first - 0.1, (first, last)
    assert last < math.log(4 * math.pi) - 0.05, last     # better than a uniform proposal
    # instance 0: plain sweeps only; instance 1: sweeps + learned rounds
    rng = np.random.default_rng(2)
    n = 600
    lat.series_reset(n, with_charge=False)
    acc = 0
    for k in range(n):
        lat.equil(2)
        ox, oy = rng.integers(0, 32, 2)
        acc += int(block_moves(lat, net, [1], int(ox), int(oy))[0])
        lat.sample(k)
    obs, _, _ = lat.get_series()
    e0, s0, _ = exact.mean_err(obs[0, 100:, 10])
    e1, s1, _ = exact.mean_err(obs[1, 100:, 10])
    assert abs(e0 - e1) < 4 * math.hypot(s0, s1), (e0, e1, s0, s1)
    print("loss %.3f -> %.3f, learned block acceptance %.3f" % (first, last, acc / (n * 4)))


@pytest.fixture(scope="module")
def learned_comparison():
    # the comparison protocol at a reduced size: L = 128, T = 0.3, 8 replicas
    # per method, 20000 swap samples of preparation, 4000 samples of 4 sweeps
    from nanomagmc_amd.learn import compare_ess
    r = compare_ess(L=128, temps=(0.3,), replicas=8, anneal_sweeps=2000, collect_rounds=1000, collect_blocks=10,
                    train_steps=300, batch=128, eq_samples=20000, nsamples=4000, timing_samples=100, log=print)
    row = r["temps"][0]
    print(r["s_per_sample"], row)
    return r, row


def test_learned_rounds_against_plain_sweeps_tau_and_ess(learned_comparison):
    # what the learned rounds do that can be measured at this size, each
    # against its error: the training loss falls; E's windows close and
    # E's gain is in the measured band (0.24 +- 0.02 for swaps, 0.33 +- 0.03
    # for vMF: a round costs more than it decorrelates E); the observables
    # agree with plain sweeps; and swaps change the topological charge more
    # often per sweep than plain sweeps (0.23 against 0.05 per 1000 sweeps)
    r, row = learned_comparison
    assert r["train_blocks"] == 10_000
    assert row["loss"][1] < row["loss"][0]
    for m in ("vmf", "swap"):
        assert 0.0 <= row["acceptance"][m] <= 1.0
        assert r["s_per_sample"][m] > r["s_per_sample"]["plain"] > 0
    assert row["acceptance"]["swap"] > 0.05
    for m in ("plain", "vmf", "swap"):
        assert row[m]["E"]["window_closed"], m
        for key in ("E", "absQ"):
            st = row[m][key]
            assert math.isfinite(st["mean"]) and st["tau_int"] >= 0.5 and st["ess_per_s"] > 0
            d = st["mean"] - row["plain"][key]["mean"]
            assert abs(d) < 4 * math.hypot(st["err"], row["plain"][key]["err"]) + 1e-9, (m, key, d)
    for m in ("vmf", "swap"):
        g, e = row["gain_E"][m], row["gain_E_err"][m]
        assert 0.08 < g < 1.0 and e < 0.3 * g, (m, g, e)
    assert row["swap"]["q_changes_per_ksweep"] > 2 * row["plain"]["q_changes_per_ksweep"]


@pytest.mark.xfail(reason="the >= 3x target for |Q| is not reached: at T <= 0.3 no method's |Q| series "
                          "decorrelates within the run (the multi-chain windows do not close), and a swap "
                          "round costs about 4x a plain sample of 4 sweeps", strict=False)
def test_learned_gain_target_for_abs_q(learned_comparison):
    _, row = learned_comparison
    for m in ("plain", "swap"):
        assert row[m]["absQ"]["window_closed"], m
    assert row["gain_absQ"]["swap"] - 2 * row["gain_absQ_err"]["swap"] >= 3.0


def stencil_slots(D):
    """J and D per stencil slot (2 = -x, 3 = +x, 4 = -y, 5 = +y) of the Neel table."""
    d, Jv, Dv = neel2d(D)
    J6, D6 = np.zeros(6), np.zeros((6, 3))
    slot = {(0, 1): 5, (0, -1): 4, (1, 0): 3, (-1, 0): 2}
    for k in range(4):
        g = slot[(int(d[k][0]), int(d[k][1]))]
        J6[g], D6[g] = Jv[k], Dv[k]
    return J6, D6


def constant_net(mu, kappa=1.0):
    """A network whose output is the same (mu, kappa) at every block site."""
    from nanomagmc_amd.learn import Net
    from oracle import netref
    p = np.zeros(netref.n_params(), dtype=np.float32)
    off = last_bias_offset()
    p[off:off + 4] = list(mu) + [math.log(math.expm1(kappa - 0.05))]
    net = Net(seed=5)
    net.set_params(p)
    return net


def test_block_swaps_match_oracle(orc):
    # a random network (so the seed sites differ between blocks) on a noisy
    # ferromagnetic lattice: the seed site, the rotated block, |A - B|^2, the
    # block charge difference, dE_s and every accept decision against the
    # FP64 oracle
    from nanomagmc_amd.learn import Net, block_swaps
    from oracle import netref
    L, D, h, T = 64, 0.3, 0.06, 0.3
    lat = lattice(L, D)
    lat.set_params(1 / T, h)
    rng = np.random.default_rng(8)
    s = np.zeros((L * L, 3))
    s[:, 2] = 1.0
    s += 0.25 * rng.normal(size=s.shape)
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    before = s.astype(np.float32).astype(np.float64)
    lat.set_spins(0, before)
    p = netref.init_params(rng).astype(np.float32)
    net = Net(seed=6)
    net.set_params(p)
    off = (3, 21)
    iters, R0 = 80, 5.0
    acc, stats, prop = block_swaps(lat, net, [0], *off, relax_iters=iters, core_radius=R0, diagnostics=True)
    after = lat.get_spins(0)
    J6, D6 = stencil_slots(D)
    d, Jv, Dv = neel2d(D)
    f = orc.Field(box_locs(2, L), 2, L, 1, d, Jv, Dv)
    H = [0, 0, h]
    E0 = f.energy(before, H, use_d=True, convention=1)
    blocks = [(32 * i + off[0], 32 * j + off[1]) for i in range(2) for j in range(2)]
    P, _ = patches_of(before, L, blocks)
    _, O = netref.forward(p.astype(np.float64), P, use_bf16=True)
    naccepted, seeds, created = 0, set(), 0
    for b, (bx, by) in enumerate(blocks):
        ring = before.reshape(L, L, 3)[np.ix_(np.arange(bx - 1, bx + 17) % L, np.arange(by - 1, by + 17) % L)]
        m = ring[0, 1:17].sum(0) + ring[17, 1:17].sum(0) + ring[1:17, 0].sum(0) + ring[1:17, 17].sum(0)
        m /= np.linalg.norm(m)
        (cu, cv), score = netref.seed_site(O[b], m)
        dev = int(stats[b, 3])
        # the device's choice is the oracle's, or ties with it to the device's fp32/bf16 rounding
        assert dev == 16 * cu + cv or score[dev // 16, dev % 16] - score[cu, cv] < 2e-3, (b, dev, (cu, cv))
        seeds.add(dev)
        xp, A, B, info = netref.swap_proposal(before, L, bx, by, O[b], J6, D6, h, iters, R0,
                                              seed=(dev // 16, dev % 16))
        assert np.abs(prop[b] - xp).max() < 2e-3, (b, np.abs(prop[b] - xp).max())
        sep = ((A - B) ** 2).sum()
        assert abs(stats[b, 1] - sep) <= 1e-2 * sep + 1e-3, (b, stats[b, 1], sep)
        assert abs(stats[b, 2] - info["dq"]) < 1e-3 and abs(info["dq"] - round(info["dq"])) < 1e-6, \
            (b, stats[b, 2], info["dq"])
        created += abs(round(info["dq"])) == 1
        idx = ((np.arange(bx, bx + 16)[:, None] % L) * L + (np.arange(by, by + 16)[None, :] % L)).reshape(-1)
        trial = before.copy()
        trial[idx] = prop[b].reshape(256, 3)
        dE = f.energy(trial, H, use_d=True, convention=1) - E0
        assert abs(stats[b, 0] - dE) <= 2e-3 * (1 + abs(dE)), (b, stats[b, 0], dE)
        w = orc.philox([b * 256, 4 << 8, 1, 0], [9, 0])
        u = np.float32(((int(w[2]) >> 8) + 0.5) / 16777216.0)
        loga = np.float32(-np.float32(1 / T) * stats[b, 0])
        accept = bool(loga >= 0 or np.log(u) < loga)
        if accept:
            assert np.array_equal(after[idx], prop[b].reshape(256, 3).astype(np.float64))
            naccepted += 1
        else:
            assert np.array_equal(after[idx], before[idx])
    assert acc[0] == naccepted
    assert len(seeds) > 1 and created >= 2


def test_block_swap_is_an_involution():
    # at negligible beta every swap is accepted; two rounds on the same
    # blocks give the lattice back
    from nanomagmc_amd.learn import block_swaps
    L, D, h = 64, 0.3, 0.06
    lat = lattice(L, D)
    lat.set_params(1e-7, h)
    lat.randomize()
    before = lat.get_spins(0)
    net = constant_net((0.0, 0.0, 1.0))
    a1 = block_swaps(lat, net, [0], 11, 5)
    mid = lat.get_spins(0)
    a2 = block_swaps(lat, net, [0], 11, 5)
    after = lat.get_spins(0)
    assert a1[0] == 4 and a2[0] == 4
    assert np.abs(mid - before).max() > 0.5
    assert np.abs(after - before).max() < 1e-5


def test_block_swaps_keep_the_physics():
    # T = 0.6: instance 0 plain sweeps, instance 1 sweeps + texture swaps
    from nanomagmc_amd.learn import block_swaps
    L, D, h, T = 64, 0.3, 0.06, 0.6
    lat = lattice(L, D, ninst=2, seed=23)
    lat.set_params(1 / T, h)
    lat.randomize()
    lat.equil(500)
    net = constant_net((0.0, 0.0, 1.0))
    rng = np.random.default_rng(4)
    n = 600
    lat.series_reset(n, with_charge=False)
    acc = 0
    for k in range(n):
        lat.equil(2)
        ox, oy = rng.integers(0, 32, 2)
        acc += int(block_swaps(lat, net, [1], int(ox), int(oy))[0])
        lat.sample(k)
    obs, _, _ = lat.get_series()
    e0, s0, _ = exact.mean_err(obs[0, 100:, 10])
    e1, s1, _ = exact.mean_err(obs[1, 100:, 10])
    print("swap acceptance %.4f, <E> plain %.2f +- %.2f, with swaps %.2f +- %.2f" % (acc / (n * 4), e0, s0, e1, s1))
    assert abs(e0 - e1) < 4 * math.hypot(s0, s1), (e0, e1, s0, s1)


@pytest.mark.parametrize("kind", ["vmf", "swap"])
def test_enqueued_rounds_equal_counted_rounds(kind):
    # rounds enqueued without outputs (no synchronisation) give the same
    # lattice as rounds that return their counts, and the running totals
    # equal the sum of the returned counts
    from nanomagmc_amd.learn import block_accepts, block_moves, block_swaps
    L, D, h, T = 64, 0.3, 0.06, 0.3
    net = constant_net((0.0, 0.0, 1.0))
    spins, totals, counted = [], [], []
    for counts in (True, False):
        lat = lattice(L, D, ninst=2, seed=31)
        lat.set_params([1 / T, 2 / T], h)
        lat.randomize()
        lat.equil(20)
        lat.reset_accepts()
        rng = np.random.default_rng(8)
        got = np.zeros(2, dtype=np.int64)
        for _ in range(6):
            ox, oy = (int(x) for x in rng.integers(0, 32, 2))
            lat.equil(1)
            if kind == "vmf":
                a = block_moves(lat, net, [1, 0], ox, oy, counts=counts)
            else:
                a = block_swaps(lat, net, [1, 0], ox, oy, counts=counts)
            if counts:
                got += np.asarray(a)[::-1]   # listed as [1, 0]
        spins.append([lat.get_spins(i) for i in range(2)])
        totals.append(block_accepts(lat, [0, 1]).astype(np.int64))
        counted.append(got)
    for a, b in zip(*spins):
        assert np.array_equal(a, b)
    assert np.array_equal(totals[0], totals[1])
    assert np.array_equal(totals[0], counted[0])
