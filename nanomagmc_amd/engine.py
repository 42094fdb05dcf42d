"""Python mirror of the reference's `state` operator API over libnmc.so.

NanoMagMC's operator boundary is the C++ class `state`
(Old/includes/state.hpp:12-52).  `Lattices` holds a batch of such states
(instances) in HBM and exposes the same operations, each applied to a list
of instances:

    state::equil(sweeps * N)   -> Lattices.equil(sweeps, inst)
    state::magnetisation()     -> Lattices.measure(...)["Mx".."Mz"]
    state::submag(0)           -> ... ["subM*"]
    state::sub4mag()           -> ... ["sub4M*"]
    state::energy()            -> ... ["E_r"]  (and the sampler's "E_s")
    state::tcharge()           -> ... q_ref (and q_wrap)
    state::change_temp/field   -> Lattices.set_params(beta, h)
    curr_state = base_state    -> Lattices.snapshot(src, dst)
    state::add_to_av           -> Lattices.accumulate_avg(i)

Only plumbing lives here: every computation runs in the HIP library.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (NMC_NOBS, NMC_TRIAL_CONE, NMC_TRIAL_SPHERE, NmcBond, NmcConfig,
                   NmcInstanceDesc, check, lib)


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ilist(inst):
    if inst is None:
        return None, 0
    a = np.ascontiguousarray(np.asarray(inst, dtype=np.int32).reshape(-1))
    return a, int(a.size)


class Bond:
    """One interaction-file record, scaled (field_type.cpp:49-71)."""

    def __init__(self, d, J, D=(0.0, 0.0, 0.0)):
        self.d = tuple(int(x) for x in d) + (0,) * (3 - len(d))
        self.J = float(J)
        self.D = tuple(float(x) for x in D)


def parse_bonds(path, dim, J_mod=1.0, D_mod=1.0):
    """The reference's interaction-file reader (field_type.cpp:39-71):
    d ints, J, 3 D components per record, operator>> semantics."""
    toks = open(path).read().split()
    bonds = []
    p = 0
    while p < len(toks):
        try:
            first = int(toks[p])
        except ValueError:
            break
        loc = [first]
        p += 1
        for _ in range(1, dim):
            loc.append(int(toks[p]))
            p += 1
        J = J_mod * float(toks[p])
        p += 1
        D = [D_mod * float(toks[p + k]) for k in range(3)]
        p += 3
        bonds.append(Bond(loc, J, D))
    return bonds


class Lattices:
    """A batch of lattice instances resident on the GPU (one nmc_ctx)."""

    def __init__(self, dim, ising, periodic, bonds, instances, use_j=True, use_d=False,
                 trial="sphere", seed=0x4E4D4300):
        L = lib()
        self.dim, self.ising, self.periodic = int(dim), bool(ising), bool(periodic)
        cfg = NmcConfig(dim=self.dim, ising=int(self.ising), periodic=int(self.periodic),
                        use_j=int(bool(use_j)), use_d=int(bool(use_d)),
                        trial=NMC_TRIAL_CONE if trial == "cone" else NMC_TRIAL_SPHERE,
                        seed=int(seed) & 0xFFFFFFFFFFFFFFFF)
        nb = len(bonds)
        barr = (NmcBond * max(nb, 1))()
        for k, b in enumerate(bonds):
            barr[k].d[:] = list(b.d)
            barr[k].J = b.J
            barr[k].D[:] = list(b.D)
        ni = len(instances)
        darr = (NmcInstanceDesc * ni)()
        self._masks = []
        for k, inst in enumerate(instances):
            n = list(inst["n"]) + [1] * (3 - len(inst["n"]))
            o = list(inst.get("origin", (0, 0, 0))) + [0] * 3
            darr[k].n[:] = n[:3]
            darr[k].origin[:] = o[:3]
            darr[k].edge = int(inst["edge"])
            m = inst.get("mask")
            if m is not None:
                m = np.ascontiguousarray(np.asarray(m, dtype=np.uint8).reshape(-1))
                self._masks.append(m)
                darr[k].mask = m.ctypes.data_as(C.POINTER(C.c_uint8))
            else:
                darr[k].mask = None
        h = C.c_void_p()
        check(L.nmc_create(C.byref(cfg), barr, nb, darr, ni, C.byref(h)))
        self._h = h
        self.ninst = ni
        self._masks = []
        ns = C.c_int()
        check(L.nmc_num_slices(self._h, C.byref(ns)))
        self.nslices = ns.value

    def close(self):
        if getattr(self, "_h", None):
            lib().nmc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        """The context's hipStream_t as an integer handle."""
        p = C.c_void_p()
        check(lib().nmc_get_stream(self._h, C.byref(p)))
        return p.value or 0

    def sync(self):
        check(lib().nmc_sync(self._h))

    def set_lds_block(self, block):
        check(lib().nmc_set_lds_block(self._h, int(block)))

    def num_sites(self, i):
        n = C.c_int64()
        check(lib().nmc_num_sites(self._h, int(i), C.byref(n)))
        return n.value

    def set_params(self, beta, h, inst=None):
        a, n = _ilist(inst)
        beta = np.ascontiguousarray(np.broadcast_to(np.asarray(beta, np.float64), (n if a is not None else self.ninst,)))
        hh = np.ascontiguousarray(np.broadcast_to(np.asarray(h, np.float64), beta.shape))
        check(lib().nmc_set_params(self._h, _iptr(a), n if a is not None else self.ninst, _dptr(beta), _dptr(hh)))

    def set_cone(self, half_angle, inst=None):
        a, n = _ilist(inst)
        cnt = n if a is not None else self.ninst
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(half_angle, np.float32), (cnt,)))
        check(lib().nmc_set_cone(self._h, _iptr(a), cnt, t.ctypes.data_as(C.POINTER(C.c_float))))

    def randomize(self, inst=None):
        a, n = _ilist(inst)
        check(lib().nmc_randomize(self._h, _iptr(a), n))

    def equil(self, sweeps, inst=None):
        a, n = _ilist(inst)
        check(lib().nmc_sweep(self._h, _iptr(a), n, int(sweeps)))

    def equil_adapt(self, sweeps, chunk=10, inst=None):
        """`sweeps` sweeps with the cone half-angles adapted on the device
        after every `chunk` sweeps; returns the final half-angles."""
        a, n = _ilist(inst)
        cnt = n if a is not None else self.ninst
        out = np.zeros(cnt, dtype=np.float32)
        check(lib().nmc_sweep_adapt(self._h, _iptr(a), n, int(sweeps), int(chunk),
                                    out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_stream_block(self, block):
        """Work-items per streaming tile: 0 (by set size), 64 or 256."""
        check(lib().nmc_set_stream_block(self._h, int(block)))

    def sweep_kinds(self, inst=None):
        """NMC_KIND_* of the kernel family sweeping each listed instance."""
        a, n = _ilist(inst)
        cnt = n if a is not None else self.ninst
        out = np.zeros(cnt, dtype=np.int32)
        check(lib().nmc_sweep_kinds(self._h, _iptr(a), n, _iptr(out)))
        return out

    def set_path(self, path):
        """'auto' (LDS-resident sweeps where the lattice fits), 'stream', or
        'cstream' (open shaped lattices streamed from compacted copies)."""
        check(lib().nmc_set_path(self._h, {"auto": 0, "stream": 1, "cstream": 2}[path]))

    def accepts(self, inst=None):
        a, n = _ilist(inst)
        cnt = n if a is not None else self.ninst
        out = np.zeros(cnt, dtype=np.uint64)
        check(lib().nmc_get_accepts(self._h, _iptr(a), n, out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def reset_accepts(self):
        check(lib().nmc_reset_accepts(self._h))

    def snapshot(self, src, dst):
        s, n = _ilist(src)
        d, m = _ilist(dst)
        assert n == m
        check(lib().nmc_snapshot(self._h, _iptr(s), _iptr(d), n))

    def measure(self, inst=None):
        """Returns (obs[n, 11], q_ref[n, nslices], q_wrap[n, nslices])."""
        a, n = _ilist(inst)
        cnt = n if a is not None else self.ninst
        obs = np.zeros((cnt, NMC_NOBS))
        qr = np.zeros((cnt, self.nslices))
        qw = np.zeros((cnt, self.nslices))
        check(lib().nmc_measure(self._h, _iptr(a), n, _dptr(obs), _dptr(qr), _dptr(qw)))
        return obs, qr, qw

    def series_reset(self, nsamples, with_charge=True):
        self.nsamples = int(nsamples)
        check(lib().nmc_series_reset(self._h, int(nsamples), int(bool(with_charge))))

    def sample(self, slot, inst=None):
        a, n = _ilist(inst)
        check(lib().nmc_sample(self._h, _iptr(a), n, int(slot)))

    def run_samples(self, nsamples, samp_sweeps, inst=None):
        """nsamples x (samp_sweeps sweeps + sample) as one replayed HIP graph."""
        a, n = _ilist(inst)
        check(lib().nmc_run_samples(self._h, _iptr(a), n, int(nsamples), int(samp_sweeps)))

    def run_samples_avg(self, nsamples, samp_sweeps, avg_inst, inst=None):
        """nsamples x (samp_sweeps sweeps + sample + add_to_av of avg_inst) as
        one replayed HIP graph."""
        a, n = _ilist(inst)
        b, m = _ilist(avg_inst)
        check(lib().nmc_run_samples_avg(self._h, _iptr(a), n, int(nsamples), int(samp_sweeps), _iptr(b), m))

    def get_series(self, inst=None):
        a, n = _ilist(inst)
        cnt = n if a is not None else self.ninst
        ns = self.nsamples
        obs = np.zeros((cnt, ns, NMC_NOBS))
        qr = np.zeros((cnt, ns, self.nslices))
        qw = np.zeros((cnt, ns, self.nslices))
        check(lib().nmc_get_series(self._h, _iptr(a), n, _dptr(obs), _dptr(qr), _dptr(qw)))
        return obs, qr, qw

    def accumulate_avg(self, i):
        check(lib().nmc_accumulate_avg(self._h, int(i)))

    def accumulate_avgs(self, inst=None):
        """add_to_av for every listed instance (all when None) in one launch."""
        a, n = _ilist(range(self.ninst) if inst is None else inst)
        check(lib().nmc_accumulate_avgs(self._h, _iptr(a), n))

    def zero_avg(self, i):
        check(lib().nmc_zero_avg(self._h, int(i)))

    def get_avg(self, i):
        out = np.zeros((self.num_sites(i), 3))
        check(lib().nmc_get_avg(self._h, int(i), _dptr(out)))
        return out

    def get_spins(self, i):
        out = np.zeros((self.num_sites(i), 3))
        check(lib().nmc_get_spins(self._h, int(i), _dptr(out)))
        return out

    def set_spins(self, i, spins):
        s = np.ascontiguousarray(np.asarray(spins, dtype=np.float64).reshape(-1, 3))
        assert s.shape[0] == self.num_sites(i)
        check(lib().nmc_set_spins(self._h, int(i), _dptr(s)))

    def delta_e(self, i, sites, trials):
        s = np.ascontiguousarray(np.asarray(sites, dtype=np.int32))
        t = np.ascontiguousarray(np.asarray(trials, dtype=np.float64).reshape(-1, 3))
        out = np.zeros(s.size)
        check(lib().nmc_delta_e(self._h, int(i), _iptr(s), _dptr(t), int(s.size), _dptr(out)))
        return out

    def replay_ising(self, i, sites, trials, k):
        s = np.ascontiguousarray(np.asarray(sites, dtype=np.int32))
        t = np.ascontiguousarray(np.asarray(trials, dtype=np.int8))
        kk = np.ascontiguousarray(np.asarray(k, dtype=np.uint32))
        acc = np.zeros(s.size, dtype=np.uint8)
        check(lib().nmc_replay_ising(self._h, int(i), _iptr(s), t.ctypes.data_as(C.POINTER(C.c_int8)),
                                     kk.ctypes.data_as(C.POINTER(C.c_uint32)), int(s.size),
                                     acc.ctypes.data_as(C.POINTER(C.c_uint8))))
        return acc


def build_weibull(dim, edge, sizes, wbeta, seed):
    """Weibull grain masks on the device (shape::weibull::check for a batch)."""
    sizes = np.ascontiguousarray(np.asarray(sizes, dtype=np.float64))
    box = edge ** dim
    out = np.zeros((sizes.size, box), dtype=np.uint8)
    check(lib().nmc_build_weibull(int(dim), int(edge), _dptr(sizes), int(sizes.size), float(wbeta),
                                  int(seed) & 0xFFFFFFFFFFFFFFFF, out.ctypes.data_as(C.POINTER(C.c_uint8))))
    return out


def crop_to_box(mask, dim, edge):
    """Tight bounding box of a box-scan mask -> (n, origin, cropped mask)."""
    m = np.asarray(mask, dtype=np.uint8).reshape((edge,) * dim)
    idx = np.nonzero(m)
    if len(idx[0]) == 0:
        raise ValueError("empty grain")
    lo = [int(a.min()) for a in idx]
    hi = [int(a.max()) + 1 for a in idx]
    sl = tuple(slice(l, h) for l, h in zip(lo, hi))
    sub = np.ascontiguousarray(m[sl])
    n = [h - l for l, h in zip(lo, hi)]
    return n, lo, sub.reshape(-1)
