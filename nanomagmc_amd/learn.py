"""Python handle on the learned block proposals of libnmc.so (nmc_net_*,
nmc_block_moves): a 4-layer 3x3 convolutional network on the matrix cores
that proposes 16 x 16 block updates of 2-D Heisenberg lattices, accepted by
exact Metropolis-Hastings.  Plumbing only; all work runs on the device."""
import ctypes as C
import math

import numpy as np

from ._lib import check, lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Net:
    def __init__(self, seed=0x4E4D4305):
        h = C.c_void_p()
        check(lib().nmc_net_create(int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(h)))
        self._h = h
        n = C.c_int()
        check(lib().nmc_net_num_params(C.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().nmc_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(self.nparams, dtype=np.float32)
        check(lib().nmc_net_get_params(self._h, _fptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(-1))
        assert p.size == self.nparams
        check(lib().nmc_net_set_params(self._h, _fptr(p)))

    def collect(self, lat, inst, nblocks, seed):
        check(lib().nmc_net_collect(self._h, lat._h, int(inst), int(nblocks), int(seed) & 0xFFFFFFFFFFFFFFFF))

    def num_samples(self):
        n = C.c_int64()
        check(lib().nmc_net_num_samples(self._h, C.byref(n)))
        return n.value

    def train(self, nsteps, batch=256, lr=1e-3, seed=1):
        loss = C.c_float()
        check(lib().nmc_net_train(self._h, int(nsteps), int(batch), float(lr), int(seed), C.byref(loss)))
        return loss.value

    def loss_grad(self, patches, targets):
        p = np.ascontiguousarray(np.asarray(patches, dtype=np.float32))
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.float32))
        n = p.shape[0]
        assert p.shape == (n, 24, 24, 4) and t.shape == (n, 256, 3)
        loss = C.c_float()
        g = np.zeros(self.nparams, dtype=np.float32)
        check(lib().nmc_net_loss_grad(self._h, _fptr(p), _fptr(t), n, C.byref(loss), _fptr(g)))
        return loss.value, g


def block_moves(lat, net, inst, off_x, off_y, diagnostics=False):
    """One round of learned block moves on each listed instance; returns the
    accepted-block counts (and per-block stats / proposals with diagnostics)."""
    a = np.ascontiguousarray(np.asarray(inst, dtype=np.int32).reshape(-1))
    acc = np.zeros(a.size, dtype=np.int64)
    stats = prop = None
    if diagnostics:
        nb = sum((lat_edge(lat, i) // 32) ** 2 for i in a)
        stats = np.zeros((nb, 3), dtype=np.float32)
        prop = np.zeros((nb, 16, 16, 3), dtype=np.float32)
    check(lib().nmc_block_moves(lat._h, net._h, a.ctypes.data_as(C.POINTER(C.c_int)), int(a.size), int(off_x),
                                int(off_y), acc.ctypes.data_as(C.POINTER(C.c_int64)),
                                _fptr(stats) if stats is not None else None,
                                _fptr(prop) if prop is not None else None))
    return (acc, stats, prop) if diagnostics else acc


def lat_edge(lat, i):
    return int(round(lat.num_sites(int(i)) ** 0.5))


def neel_bonds(D):
    """Square-lattice nearest-neighbour bonds, J = 1, interfacial DMI
    D_ij = D (z x r_ij)."""
    from .engine import Bond
    out = []
    for d in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        Dv = (-D * d[1], D * d[0], 0.0)
        out.append(Bond(d, 1.0, Dv))
    return out


def compare_ess(L=512, temps=(0.1, 0.2, 0.3), h=0.06, D=0.3, replicas=2, anneal_sweeps=3000,
                collect_rounds=100, collect_blocks=1000, train_steps=2000, batch=256, lr=1e-3,
                nsamples=2000, sweeps_per_round=4, timing_samples=100, seed=0x4E4D4305, log=None):
    """Plain checkerboard sweeps against sweeps interleaved with learned block
    rounds (sweeps_per_round : 1), on the 2-D skyrmion Hamiltonian (J = 1,
    Neel DMI D, field h along z, periodic L x L).

    Per temperature: 2 * replicas instances (half plain, half learned) in
    one context; all are annealed from random spins (T = 1 -> T over
    anneal_sweeps); one network per temperature is trained on
    collect_rounds x collect_blocks blocks of the plain instances; then
    nsamples samples of E_s and |Q| (periodic plaquette sum) are taken,
    each after sweeps_per_round sweeps (plus one block round for the
    learned instances).  The cost of one sample of each method is timed
    on its own.  Returns a dict with tau_int of E and |Q| (in samples),
    acceptance, seconds per sample and tau_int-normalised effective
    samples per second for both methods."""
    import time

    from .engine import Lattices
    from .stats import ess, mean_err
    say = log or (lambda *a: None)
    rng = np.random.default_rng(seed & 0xFFFFFFFF)
    nT = len(temps)
    R = int(replicas)
    insts = [{"n": (L, L), "edge": L}] * (nT * 2 * R)
    lat = Lattices(2, False, True, neel_bonds(D), insts, use_j=True, use_d=True, seed=seed)
    plain = [[t * 2 * R + r for r in range(R)] for t in range(nT)]
    learned = [[t * 2 * R + R + r for r in range(R)] for t in range(nT)]
    target = np.repeat(np.asarray(temps, dtype=np.float64), 2 * R)
    lat.randomize()
    t0 = time.perf_counter()
    nstep = 20
    for k in range(nstep):   # anneal from T = 1 down to each target
        T = 1.0 + (target - 1.0) * (k + 1) / nstep
        lat.set_params(list(1.0 / T), h)
        lat.equil(max(1, anneal_sweeps // nstep))
    lat.sync()
    say("annealed %d sweeps in %.1f s" % (anneal_sweeps, time.perf_counter() - t0))
    nets, losses = [], []
    t0 = time.perf_counter()
    for t in range(nT):
        net = Net(seed=seed + 17 * t)
        for k in range(collect_rounds):
            lat.equil(10, inst=plain[t])
            net.collect(lat, plain[t][k % R], collect_blocks, seed=int(rng.integers(1 << 62)))
        first = net.train(1, batch=batch, lr=0.0, seed=1)
        last = net.train(train_steps, batch=batch, lr=lr, seed=2)
        nets.append(net)
        losses.append((first, last))
    lat.sync()
    train_s = time.perf_counter() - t0
    say("trained %d networks in %.1f s: %s" % (nT, train_s, losses))

    def block_round(which):
        acc = np.zeros(nT, dtype=np.int64)
        for t in range(nT):
            ox, oy = rng.integers(0, 32, 2)
            acc[t] = int(np.sum(block_moves(lat, nets[t], which[t], int(ox), int(oy))))
        return acc

    flat_plain = [i for g in plain for i in g]
    flat_learned = [i for g in learned for i in g]
    # cost of one sample of each method, each on its own instances
    lat.sync()
    t0 = time.perf_counter()
    for _ in range(timing_samples):
        lat.equil(sweeps_per_round, inst=flat_plain)
    lat.sync()
    t_plain = (time.perf_counter() - t0) / timing_samples
    t0 = time.perf_counter()
    for _ in range(timing_samples):
        lat.equil(sweeps_per_round, inst=flat_learned)
        block_round(learned)
    lat.sync()
    t_learned = (time.perf_counter() - t0) / timing_samples
    # production run
    lat.series_reset(nsamples, with_charge=True)
    nacc = np.zeros(nT, dtype=np.int64)
    t0 = time.perf_counter()
    for k in range(nsamples):
        lat.equil(sweeps_per_round)
        nacc += block_round(learned)
        lat.sample(k)
    lat.sync()
    say("sampled %d x %d sweeps in %.1f s" % (nsamples, sweeps_per_round, time.perf_counter() - t0))
    obs, _, qw = lat.get_series()
    blocks_per_round = (L // 32) ** 2
    out = {"L": L, "h": h, "D": D, "replicas": R, "nsamples": nsamples, "sweeps_per_round": sweeps_per_round,
           "train_blocks": collect_rounds * collect_blocks, "train_steps": train_steps, "batch": batch,
           "train_s": train_s, "s_per_sample": {"plain": t_plain, "learned": t_learned}, "temps": []}
    for t, T in enumerate(temps):
        row = {"T": T, "loss": losses[t],
               "acceptance": float(nacc[t]) / (nsamples * blocks_per_round * R)}
        for name, ids in (("plain", plain[t]), ("learned", learned[t])):
            st = {}
            for key, series in (("E", obs[:, :, 10]), ("absQ", np.abs(qw[:, :, 0]))):
                taus, means, errs, closed = [], [], [], True
                for i in ids:
                    _, tau, ok = ess(series[i])
                    m, e, _ = mean_err(series[i])
                    taus.append(tau)
                    means.append(m)
                    errs.append(e)
                    closed = closed and ok
                tau = float(np.mean(taus))
                cost = t_plain if name == "plain" else t_learned
                st[key] = {"mean": float(np.mean(means)), "err": float(math.sqrt(np.sum(np.square(errs))) / len(ids)),
                           "tau_int": tau, "window_closed": closed,
                           # effective samples per second over this method's instances of this temperature,
                           # charged the method's full per-sample cost share
                           "ess_per_s": len(ids) / (2.0 * tau * cost / nT)}
            row[name] = st
        row["gain_absQ"] = row["learned"]["absQ"]["ess_per_s"] / row["plain"]["absQ"]["ess_per_s"]
        row["gain_E"] = row["learned"]["E"]["ess_per_s"] / row["plain"]["E"]["ess_per_s"]
        out["temps"].append(row)
    return out
