"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the nanomagmc_amd package.  Each function restates a piece of the
reference (file:line in oracle.cpp).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(HERE, "liboracle.so")
_ROC = os.path.join(HERE, "librocphilox.so")
_lib = None
_roc = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        P, i, d, u32 = C.POINTER, C.c_int, C.c_double, C.c_uint32
        i64 = C.c_int64
        L.orc_philox.argtypes = [P(u32), P(u32), P(u32)]
        L.orc_parse_bonds.argtypes = [C.c_char_p, i, d, d, P(i), P(d), P(d), i]
        L.orc_parse_bonds.restype = i
        L.orc_weibull_r0.argtypes = [d]
        L.orc_weibull_r0.restype = d
        L.orc_weibull_check.argtypes = [P(i), i, i, d, d, d]
        L.orc_weibull_check.restype = i
        L.orc_build.argtypes = [i, i, C.c_char, d, d, P(d), i64, i, P(i), i64]
        L.orc_build.restype = i
        L.orc_set_neigh.argtypes = [i, P(i), i, i, i, i, P(i), P(i), P(i), P(i), P(i), i]
        L.orc_energy.argtypes = [i, P(d), P(i), P(i), P(i), i, P(d), P(d), P(d), i, i, i]
        L.orc_energy.restype = d
        L.orc_delta_e.argtypes = [i, P(d), P(d), P(i), P(i), P(i), i, P(d), P(d), P(d), i, i]
        L.orc_delta_e.restype = d
        L.orc_mag.argtypes = [i, P(d), P(d)]
        L.orc_submag.argtypes = [i, P(d), P(i), i, i, P(d)]
        L.orc_sub4mag.argtypes = [i, P(d), P(i), i, i, P(d)]
        L.orc_solid_angle.argtypes = [P(d), P(d), P(d)]
        L.orc_solid_angle.restype = d
        L.orc_tcharge.argtypes = [i, P(d), P(i), i, i, P(i), P(d)]
        L.orc_atoln.argtypes = [d, d, P(d), P(d)]
        L.orc_norm.argtypes = [P(d), i]
        L.orc_norm.restype = d
        L.orc_equil_replay.argtypes = [i, P(d), P(i), P(i), P(i), i, P(d), P(d), P(d), i, i, d,
                                       i64, P(i), P(d), P(d), P(C.c_uint8)]
        L.orc_checkerboard_ising.argtypes = [i, P(d), P(i), P(i), P(i), i, P(d), P(d), P(d), i, i, d,
                                             P(i64), P(i), C.c_uint64, u32, C.c_uint64, i]
        L.orc_accept_threshold.argtypes = [d, d]
        L.orc_accept_threshold.restype = u32
        L.orc_bench_reference.argtypes = [i, P(i), P(i), P(i), i, P(d), P(d), i, i, i, d, d, i, i64,
                                          u32, P(d)]
        L.orc_bench_reference.restype = d
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def rocrand_philox(ctr, key):
    """rocRAND's Philox4x32-10 block function (installed header, host path)."""
    global _roc
    if _roc is None:
        if not os.path.exists(_ROC):
            build()
        _roc = C.CDLL(_ROC)
        _roc.rocphilox.argtypes = [C.POINTER(C.c_uint)] * 3
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    _roc.rocphilox(_p(c, C.c_uint), _p(k, C.c_uint), _p(o, C.c_uint))
    return o


def parse_bonds(path, dim, J_mod=1.0, D_mod=1.0, max_bonds=4096):
    d = np.zeros((max_bonds, 3), dtype=np.int32)
    J = np.zeros(max_bonds)
    D = np.zeros((max_bonds, 3))
    n = lib().orc_parse_bonds(path.encode(), dim, J_mod, D_mod, _p(d, C.c_int), _p(J, C.c_double),
                              _p(D, C.c_double), max_bonds)
    if n < 0:
        raise FileNotFoundError(path)
    return d[:n].copy(), J[:n].copy(), D[:n].copy()


def build_sites(dim, edge, shape, size=1.0, wbeta=1.0, uniforms=None, faithful_carry=False):
    total = edge ** dim
    u = np.ascontiguousarray(np.zeros(1) if uniforms is None else np.asarray(uniforms, dtype=np.float64))
    locs = np.zeros((total + 1, 3), dtype=np.int32)
    n = lib().orc_build(dim, edge, shape.encode(), float(size), float(wbeta), _p(u, C.c_double), u.size,
                        int(faithful_carry), _p(locs, C.c_int), total + 1)
    if n < 0:
        raise ValueError("orc_build failed")
    return locs[:n].copy()


class Field:
    """A reference field_type: sites, neighbour lists, interaction table."""

    def __init__(self, locs, dim, edge, periodic, bd, J, D):
        self.locs = np.ascontiguousarray(locs, dtype=np.int32)
        self.n = len(self.locs)
        self.dim, self.edge, self.periodic = dim, edge, int(periodic)
        self.bd = np.ascontiguousarray(bd, dtype=np.int32)
        self.J = np.ascontiguousarray(J, dtype=np.float64)
        self.D = np.ascontiguousarray(D, dtype=np.float64)
        self.nb = len(self.J)
        self.neigh = np.zeros((self.n, max(self.nb, 1)), dtype=np.int32)
        self.nchoice = np.zeros_like(self.neigh)
        self.ncount = np.zeros(self.n, dtype=np.int32)
        self.adj = np.zeros((self.n, 4), dtype=np.int32)
        self.adj_wrap = np.zeros((self.n, 4), dtype=np.int32)
        L = lib()
        for wrap, adj in ((0, self.adj), (1, self.adj_wrap)):
            L.orc_set_neigh(self.n, _p(self.locs, C.c_int), dim, edge, self.periodic, self.nb,
                            _p(self.bd, C.c_int), _p(self.neigh, C.c_int), _p(self.nchoice, C.c_int),
                            _p(self.ncount, C.c_int), _p(adj, C.c_int), wrap)

    def _args(self):
        return (_p(self.neigh, C.c_int), _p(self.nchoice, C.c_int), _p(self.ncount, C.c_int), self.nb,
                _p(self.J, C.c_double), _p(self.D, C.c_double))

    def energy(self, s, H, use_j=True, use_d=False, convention=0):
        s = np.ascontiguousarray(s, dtype=np.float64)
        H = np.ascontiguousarray(H, dtype=np.float64)
        return lib().orc_energy(self.n, _p(s, C.c_double), *self._args(), _p(H, C.c_double), int(use_j),
                                int(use_d), int(convention))

    def delta_e(self, s, pos, trial, H, use_j=True, use_d=False):
        s = np.ascontiguousarray(s, dtype=np.float64)
        t = np.ascontiguousarray(trial, dtype=np.float64)
        H = np.ascontiguousarray(H, dtype=np.float64)
        return lib().orc_delta_e(int(pos), _p(t, C.c_double), _p(s, C.c_double), *self._args(),
                                 _p(H, C.c_double), int(use_j), int(use_d))

    def mag(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        o = np.zeros(3)
        lib().orc_mag(self.n, _p(s, C.c_double), _p(o, C.c_double))
        return o

    def submag(self, s, subnumber=0):
        s = np.ascontiguousarray(s, dtype=np.float64)
        o = np.zeros(3)
        lib().orc_submag(self.n, _p(s, C.c_double), _p(self.locs, C.c_int), self.dim, subnumber,
                         _p(o, C.c_double))
        return o

    def sub4mag(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        o = np.zeros(3)
        lib().orc_sub4mag(self.n, _p(s, C.c_double), _p(self.locs, C.c_int), self.dim, self.edge,
                          _p(o, C.c_double))
        return o

    def tcharge(self, s, wrap=False):
        s = np.ascontiguousarray(s, dtype=np.float64)
        o = np.zeros(self.edge if self.dim == 3 else 1)
        adj = self.adj_wrap if wrap else self.adj
        lib().orc_tcharge(self.n, _p(s, C.c_double), _p(self.locs, C.c_int), self.dim, self.edge,
                          _p(adj, C.c_int), _p(o, C.c_double))
        return o

    def equil_replay(self, s, H, beta, sites, trials, u, use_j=True, use_d=False):
        """Mutates s in place; returns accept flags (state.cpp:180-202)."""
        sites = np.ascontiguousarray(sites, dtype=np.int32)
        trials = np.ascontiguousarray(trials, dtype=np.float64).reshape(-1, 3)
        u = np.ascontiguousarray(u, dtype=np.float64)
        H = np.ascontiguousarray(H, dtype=np.float64)
        acc = np.zeros(sites.size, dtype=np.uint8)
        lib().orc_equil_replay(self.n, _p(s, C.c_double), *self._args(), _p(H, C.c_double), int(use_j),
                               int(use_d), float(beta), sites.size, _p(sites, C.c_int),
                               _p(trials, C.c_double), _p(u, C.c_double), _p(acc, C.c_uint8))
        return acc

    def colour_index(self, n_box, origin=(0, 0, 0)):
        """(element, colour) of each site in the device's colour-split layout
        of an instance box with reference extents n_box placed at origin."""
        loc = self.locs - np.asarray(origin, dtype=np.int32)[None, :]
        if self.dim == 3:
            i0, i1, i2 = loc[:, 0], loc[:, 1], loc[:, 2]
            a1, a2 = n_box[1], n_box[2]
        else:
            i0 = np.zeros(self.n, dtype=np.int64)
            i1, i2 = loc[:, 0], loc[:, 1]
            a1, a2 = n_box[0], n_box[1]
        # periodic boxes keep their (even) extent; open boxes are padded to a
        # multiple of 8 along the fast axis
        a2 = a2 if self.periodic else (a2 + 7) // 8 * 8
        half = a2 // 2
        e = ((i0.astype(np.int64) * a1 + i1) * half + (i2 >> 1)).astype(np.int64)
        c = ((i0 + i1 + i2) & 1).astype(np.int32)
        return e, c

    def checkerboard_ising(self, s, H, beta, e, c, seed, inst, sweep0, nsweeps, use_j=True, use_d=False):
        e = np.ascontiguousarray(e, dtype=np.int64)
        c = np.ascontiguousarray(c, dtype=np.int32)
        H = np.ascontiguousarray(H, dtype=np.float64)
        lib().orc_checkerboard_ising(self.n, _p(s, C.c_double), *self._args(), _p(H, C.c_double), int(use_j),
                                     int(use_d), float(beta), _p(e, C.c_int64), _p(c, C.c_int), int(seed),
                                     int(inst), int(sweep0), int(nsweeps))

    def bench_reference(self, ising, h, beta, nthreads, attempts, use_j=True, use_d=False, seed=1):
        secs = C.c_double()
        rate = lib().orc_bench_reference(self.n, *self._args(), int(ising), int(use_j), int(use_d), float(h),
                                         float(beta), int(nthreads), int(attempts), int(seed), C.byref(secs))
        return rate, secs.value


def bench_lib(out_dir):
    """The restatement built with `make benchref` into out_dir (the timed CPU
    baseline, compiled for the host it runs on); returns its
    orc_bench_reference."""
    subprocess.run(["make", "-s", "-C", HERE, "benchref", "OUT=" + os.path.abspath(out_dir)], check=True)
    L = C.CDLL(os.path.join(os.path.abspath(out_dir), "libbenchref.so"))
    P, i, d, u32, i64 = C.POINTER, C.c_int, C.c_double, C.c_uint32, C.c_int64
    f = L.orc_bench_reference
    f.argtypes = [i, P(i), P(i), P(i), i, P(d), P(d), i, i, i, d, d, i, i64, u32, P(d)]
    f.restype = d
    return f


def bench_reference(fn, field, ising, h, beta, nthreads, attempts, use_j=True, use_d=False, seed=1):
    """Field.bench_reference through another build of the restatement."""
    secs = C.c_double()
    rate = fn(field.n, *field._args(), int(ising), int(use_j), int(use_d), float(h), float(beta), int(nthreads),
              int(attempts), int(seed), C.byref(secs))
    return rate, secs.value


def accept_threshold(dE, beta):
    return lib().orc_accept_threshold(float(dE), float(beta))


def solid_angle(s1, s2, s3):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (s1, s2, s3)]
    return lib().orc_solid_angle(*[_p(x, C.c_double) for x in a])


def atoln(amean, asd):
    m, s = C.c_double(), C.c_double()
    lib().orc_atoln(float(amean), float(asd), C.byref(m), C.byref(s))
    return m.value, s.value


def weibull_r0(wbeta):
    return lib().orc_weibull_r0(float(wbeta))


def nn_bonds(dim, J=1.0, D=None):
    """Nearest-neighbour table in the reference file's record order
    (Old/Js/3d_nn_bloch.dat: +z, -z, +y, -y, +x, -x for 3-D)."""
    if dim == 3:
        d = np.array([[0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0]], dtype=np.int32)
    else:
        d = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0]], dtype=np.int32)
    Jv = np.full(len(d), float(J))
    Dv = np.zeros((len(d), 3)) if D is None else np.asarray(D, dtype=np.float64)
    return d, Jv, Dv
