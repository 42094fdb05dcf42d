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


def block_moves(lat, net, inst, off_x, off_y, diagnostics=False, counts=True):
    """One round of learned block moves on each listed instance; returns the
    accepted-block counts (and per-block stats / proposals with diagnostics).
    With counts=False (and no diagnostics) the round is only enqueued and
    None is returned; block_accepts() reads the running totals."""
    a = np.ascontiguousarray(np.asarray(inst, dtype=np.int32).reshape(-1))
    acc = np.zeros(a.size, dtype=np.int64) if (counts or diagnostics) else None
    stats = prop = None
    if diagnostics:
        nb = sum((lat_edge(lat, i) // 32) ** 2 for i in a)
        stats = np.zeros((nb, 4), dtype=np.float32)
        prop = np.zeros((nb, 16, 16, 3), dtype=np.float32)
    check(lib().nmc_block_moves(lat._h, net._h, a.ctypes.data_as(C.POINTER(C.c_int)), int(a.size), int(off_x),
                                int(off_y), acc.ctypes.data_as(C.POINTER(C.c_int64)) if acc is not None else None,
                                _fptr(stats) if stats is not None else None,
                                _fptr(prop) if prop is not None else None))
    return (acc, stats, prop) if diagnostics else acc


def block_swaps(lat, net, inst, off_x, off_y, relax_iters=40, core_radius=5.0, diagnostics=False, counts=True):
    """One round of learned texture swaps (create or remove a skyrmion per
    block, at the site the network picks, by an exact measure-preserving
    involution); returns the accepted-block counts (and per-block
    (dE, |A - B|^2, Q(B) - Q(A), seed site u * 16 + v) / rotated blocks with
    diagnostics), or None with counts=False (the round is only enqueued;
    block_accepts() reads the running totals)."""
    a = np.ascontiguousarray(np.asarray(inst, dtype=np.int32).reshape(-1))
    acc = np.zeros(a.size, dtype=np.int64) if (counts or diagnostics) else None
    stats = prop = None
    if diagnostics:
        nb = sum((lat_edge(lat, i) // 32) ** 2 for i in a)
        stats = np.zeros((nb, 4), dtype=np.float32)
        prop = np.zeros((nb, 16, 16, 3), dtype=np.float32)
    check(lib().nmc_block_swaps(lat._h, net._h, a.ctypes.data_as(C.POINTER(C.c_int)), int(a.size), int(off_x),
                                int(off_y), int(relax_iters), float(core_radius),
                                acc.ctypes.data_as(C.POINTER(C.c_int64)) if acc is not None else None,
                                _fptr(stats) if stats is not None else None,
                                _fptr(prop) if prop is not None else None))
    return (acc, stats, prop) if diagnostics else acc


def block_accepts(lat, inst):
    """Accepted learned blocks per listed instance since the context was made
    or its accept counters were reset."""
    a = np.ascontiguousarray(np.asarray(inst, dtype=np.int32).reshape(-1))
    out = np.zeros(a.size, dtype=np.uint64)
    check(lib().nmc_get_block_accepts(lat._h, a.ctypes.data_as(C.POINTER(C.c_int)), int(a.size),
                                      out.ctypes.data_as(C.POINTER(C.c_uint64))))
    return out


def lat_edge(lat, i):
    """Edge of a learned-move lattice (periodic 2-D, so the full edge x edge box)."""
    n = lat.num_sites(int(i))
    edge = int(round(n ** 0.5))
    assert edge * edge == n, "learned block moves need square periodic lattices"
    return edge


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
                collect_rounds=10000, collect_blocks=10, train_steps=2000, batch=256, lr=1e-3,
                eq_samples=20000, nsamples=4000, sweeps_per_round=4, timing_samples=100, methods=("vmf", "swap"),
                relax_iters=40, core_radius=5.0, seed=0x4E4D4305, log=None):
    """Plain checkerboard sweeps against sweeps interleaved with learned block
    rounds (sweeps_per_round : 1), on the 2-D skyrmion Hamiltonian (J = 1,
    Neel DMI D, field h along z, periodic L x L).  Learned methods: "vmf"
    (block proposals drawn from one network per temperature's von
    Mises-Fisher laws) and "swap" (texture swaps whose skyrmion seed site one
    network, trained on every temperature's blocks, picks; one round call
    covers every swap instance).

    Protocol, per temperature (1 + len(methods)) * replicas instances in one
    context:
    1. all instances are annealed from random spins (T = 1 -> T over
       anneal_sweeps plain sweeps);
    2. the networks are trained on collect_blocks blocks from each of
       collect_rounds configurations of the plain instances, 10 plain sweeps
       apart (the plain sampler's configurations);
    3. every instance is brought towards equilibrium by eq_samples samples of
       the swap sampler, the faster of the two at moving the topological
       charge, so that every method starts stationary (any valid sampler may
       prepare the start; each chain keeps its own state);
    4. each method then runs nsamples samples of its own dynamics (E_s and
       |Q|, the periodic plaquette sum, after every sweeps_per_round sweeps
       plus, for learned methods, one block round).
    The cost of one sample of each method is timed on its own instances.
    Returns tau_int of E and |Q| (in samples, over the replicas together with
    the multi-chain autocorrelation of stats.tau_int_chains, which charges
    chains stuck at different values; with whether the window closed and the
    mean per-chain tau_int beside it), acceptance, seconds per sample,
    tau_int-normalised effective samples per second and the gains over plain
    with their standard errors (Sokal's var(tau) = 2 (2 W + 1) tau^2 / n,
    n = replicas x samples).  A window that does not close makes tau_int a
    lower bound; for plain that makes the gain a lower bound too."""
    import time

    from .engine import Lattices
    from .stats import mean_err, tau_int_chains
    say = log or (lambda *a: None)
    rng = np.random.default_rng(seed & 0xFFFFFFFF)
    nT = len(temps)
    R = int(replicas)
    names = ["plain"] + list(methods)
    G = len(names)
    insts = [{"n": (L, L), "edge": L}] * (nT * G * R)
    lat = Lattices(2, False, True, neel_bonds(D), insts, use_j=True, use_d=True, seed=seed)
    group = {nm: [[t * G * R + gi * R + r for r in range(R)] for t in range(nT)] for gi, nm in enumerate(names)}
    every = list(range(nT * G * R))
    target = np.repeat(np.asarray(temps, dtype=np.float64), G * R)
    lat.randomize()
    t0 = time.perf_counter()
    nstep = 20
    for k in range(nstep):   # anneal from T = 1 down to each target
        T = 1.0 + (target - 1.0) * (k + 1) / nstep
        lat.set_params(list(1.0 / T), h)
        lat.equil(max(1, anneal_sweeps // nstep))
    lat.sync()
    say("annealed %d sweeps in %.1f s" % (anneal_sweeps, time.perf_counter() - t0))
    t0 = time.perf_counter()
    plain = group["plain"]
    nets, losses = [], []
    if "vmf" in methods:
        for t in range(nT):
            net = Net(seed=seed + 17 * t)
            for k in range(collect_rounds):
                lat.equil(10, inst=plain[t])
                net.collect(lat, plain[t][k % R], collect_blocks, seed=int(rng.integers(1 << 62)))
            first = net.train(1, batch=batch, lr=0.0, seed=1)
            last = net.train(train_steps, batch=batch, lr=lr, seed=2)
            nets.append(net)
            losses.append((first, last))
    swap_net, swap_loss = None, None
    if "swap" in methods:
        swap_net = Net(seed=seed + 101)
        for k in range(collect_rounds):
            t = k % nT
            lat.equil(10, inst=plain[t])
            swap_net.collect(lat, plain[t][(k // nT) % R], collect_blocks, seed=int(rng.integers(1 << 62)))
        first = swap_net.train(1, batch=batch, lr=0.0, seed=1)
        swap_loss = (first, swap_net.train(train_steps, batch=batch, lr=lr, seed=2))
    lat.sync()
    train_s = time.perf_counter() - t0
    say("trained networks in %.1f s: vmf %s swap %s" % (train_s, losses, swap_loss))

    swap_ids = [i for g in group.get("swap", []) for i in g]

    def block_round(nm, ids=None):
        # enqueued only: acceptance is read from the running totals at the end
        if nm == "vmf":
            for t in range(nT):
                ox, oy = (int(x) for x in rng.integers(0, 32, 2))
                block_moves(lat, nets[t], group[nm][t], ox, oy, counts=False)
        else:
            ox, oy = (int(x) for x in rng.integers(0, 32, 2))
            block_swaps(lat, swap_net, ids if ids is not None else swap_ids, ox, oy, relax_iters, core_radius,
                        counts=False)

    if eq_samples and swap_net is not None:
        t0 = time.perf_counter()
        for k in range(eq_samples):
            lat.equil(sweeps_per_round)
            block_round("swap", every)
        lat.sync()
        say("equilibrated every instance with %d swap samples in %.1f s" % (eq_samples, time.perf_counter() - t0))
    # cost of one sample of each method, each on its own instances
    cost = {}
    for nm in names:
        ids = [i for g in group[nm] for i in g]
        lat.sync()
        t0 = time.perf_counter()
        for _ in range(timing_samples):
            lat.equil(sweeps_per_round, inst=ids)
            if nm != "plain":
                block_round(nm)
        lat.sync()
        cost[nm] = (time.perf_counter() - t0) / timing_samples
    # production run
    lat.series_reset(nsamples, with_charge=True)
    lat.reset_accepts()
    t0 = time.perf_counter()
    for k in range(nsamples):
        lat.equil(sweeps_per_round)
        for nm in methods:
            block_round(nm)
        lat.sample(k)
    lat.sync()
    nacc = {nm: np.array([int(np.sum(block_accepts(lat, group[nm][t]))) for t in range(nT)]) for nm in methods}
    say("sampled %d x %d sweeps in %.1f s" % (nsamples, sweeps_per_round, time.perf_counter() - t0))
    obs, _, qw = lat.get_series()
    blocks_per_round = (L // 32) ** 2
    out = {"L": L, "h": h, "D": D, "replicas": R, "nsamples": nsamples, "eq_samples": eq_samples,
           "sweeps_per_round": sweeps_per_round, "train_blocks": collect_rounds * collect_blocks,
           "train_configurations": collect_rounds,
           "train_steps": train_steps, "batch": batch, "train_s": train_s, "relax_iters": relax_iters,
           "core_radius": core_radius, "s_per_sample": cost, "swap_loss": swap_loss, "temps": []}
    for t, T in enumerate(temps):
        row = {"T": T, "loss": losses[t] if losses else swap_loss,
               "acceptance": {nm: float(nacc[nm][t]) / (nsamples * blocks_per_round * R) for nm in methods}}
        for nm in names:
            st = {}
            for key, series in (("E", obs[:, :, 10]), ("absQ", np.abs(qw[:, :, 0]))):
                X = series[group[nm][t]]
                tau, W, closed = tau_int_chains(X)
                means, errs, chain_taus = [], [], []
                for x in X:
                    m, _, tc = mean_err(x)
                    means.append(m)
                    chain_taus.append(tc)
                # the mean's error from the multi-chain tau and the pooled variance
                var = X.var(ddof=1) if X.size > 1 else 0.0
                st[key] = {"mean": float(np.mean(means)), "err": float(math.sqrt(2.0 * tau * var / X.size)),
                           "tau_int": tau, "tau_rel_err": float(math.sqrt(2.0 * (2 * W + 1) / X.size)),
                           "window_closed": bool(closed), "tau_int_per_chain": float(np.mean(chain_taus)),
                           # effective samples per second over this method's instances of this temperature,
                           # charged the method's full per-sample cost share
                           "ess_per_s": R / (2.0 * tau * cost[nm] / nT)}
            # topological changes: samples whose (integer) charge differs from the previous sample
            q = np.rint(qw[group[nm][t], :, 0])
            st["q_changes_per_ksweep"] = float((q[:, 1:] != q[:, :-1]).sum()) / (R * (nsamples - 1) *
                                                                                  sweeps_per_round) * 1000.0
            row[nm] = st
        for key in ("absQ", "E"):
            g, ge = {}, {}
            for nm in methods:
                r = row[nm][key]["ess_per_s"] / row["plain"][key]["ess_per_s"]
                g[nm] = r
                ge[nm] = r * math.hypot(row[nm][key]["tau_rel_err"], row["plain"][key]["tau_rel_err"])
            row["gain_" + key] = g
            row["gain_%s_err" % key] = ge
        out["temps"].append(row)
    return out
