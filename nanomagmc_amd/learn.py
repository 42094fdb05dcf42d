"""Python handle on the learned block proposals of libnmc.so (nmc_net_*,
nmc_block_moves): a 4-layer 3x3 convolutional network on the matrix cores
that proposes 16 x 16 block updates of 2-D Heisenberg lattices, accepted by
exact Metropolis-Hastings.  Plumbing only; all work runs on the device."""
import ctypes as C

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
