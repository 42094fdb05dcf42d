"""nmc_measure_noalloc (the measurement with its partial sums in the
library's static device scratch) against nmc_measure, through the C ABI:
the same observables bit for bit on two calls in a row, the second with a
smaller instance list, and the errors for an unprepared list and for a list
whose partial sums exceed the scratch.  Needs an MI355X."""
import ctypes as C

import numpy as np
import pytest

from tests.fixtures import nn2d
from tests.gpu_util import engine_bonds

pytestmark = pytest.mark.gpu


def noalloc(lat, inst, charges=True):
    from nanomagmc_amd._lib import lib
    a = None if inst is None else np.ascontiguousarray(np.asarray(inst, dtype=np.int32))
    cnt = lat.ninst if a is None else a.size
    obs = np.full((cnt, 11), np.nan)
    qr = np.full((cnt, lat.nslices), np.nan)
    qw = np.full((cnt, lat.nslices), np.nan)
    dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))   # noqa: E731
    rc = lib().nmc_measure_noalloc(lat._h, None if a is None else a.ctypes.data_as(C.POINTER(C.c_int)),
                                   0 if a is None else int(a.size), dp(obs), dp(qr) if charges else None,
                                   dp(qw) if charges else None)
    return rc, obs, qr, qw


def grains_lattice(orc, bloch3d):
    from nanomagmc_amd import Lattices
    rng = np.random.default_rng(17)
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    insts = []
    for n in (10, 7, 12, 9, 14, 8):
        m = (rng.random(n ** 3) < 0.7).astype(np.uint8)
        insts.append({"n": (n,) * 3, "edge": n, "mask": m})
    lat = Lattices(3, False, False, engine_bonds(bd, J, D), insts, use_d=True, seed=23)
    lat.set_params(np.linspace(0.5, 3.0, 6), [0.0, 0.05, 0.1, 0.0, 0.02, 0.07])
    lat.randomize()
    lat.equil(5)
    return lat


def test_noalloc_measure_equals_measure_twice_shrinking(orc, bloch3d):
    lat = grains_lattice(orc, bloch3d)
    want = lat.measure()
    rc, obs, qr, qw = noalloc(lat, None)
    assert rc == 0
    assert np.array_equal(obs, want[0]) and np.array_equal(qr, want[1]) and np.array_equal(qw, want[2])
    lat.equil(3)
    small = [4, 1]
    want = lat.measure(small)          # prepares the list (and is the reference)
    rc, obs, qr, qw = noalloc(lat, small)
    assert rc == 0
    assert np.array_equal(obs, want[0]) and np.array_equal(qr, want[1]) and np.array_equal(qw, want[2])
    # without charges the same sums (the charge pass is skipped)
    rc, obs2, _, _ = noalloc(lat, small, charges=False)
    assert rc == 0 and np.array_equal(obs2, want[0])


def test_noalloc_measure_ising_periodic():
    from nanomagmc_amd import Lattices
    d, J, D = nn2d()
    lat = Lattices(2, True, True, engine_bonds(d, J, D), [{"n": (32, 32), "edge": 32}] * 5, seed=29)
    lat.set_params(np.linspace(0.3, 0.6, 5), 0.05)
    lat.randomize()
    lat.equil(4)
    for inst in (None, [3, 0, 2], [1]):
        want = lat.measure(inst)
        rc, obs, qr, qw = noalloc(lat, inst)
        assert rc == 0 and np.array_equal(obs, want[0])
        assert np.all(qr == 0) and np.all(qw == 0)


def test_noalloc_measure_rejects_unprepared_and_oversize_lists():
    from nanomagmc_amd import Lattices
    from nanomagmc_amd._lib import lib
    d3 = [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)]
    bonds = engine_bonds(d3, [1.0] * 6, [(0.0, 0.0, 0.0)] * 6)
    # 32 open boxes of 16 x 16 x 512: 64 measurement tiles each with 15 + 2 * 512
    # partial sums per tile, 17 MB in all, more than the 16 MiB scratch
    lat = Lattices(3, False, False, bonds, [{"n": (16, 16, 512), "edge": 512}] * 32, seed=31)
    lat.randomize()
    rc, _, _, _ = noalloc(lat, [5, 6])
    assert rc == -1 and "work lists" in lib().nmc_last_error().decode()
    rc, _, _, _ = noalloc(lat, None)
    assert rc == -1 and "scratch" in lib().nmc_last_error().decode()
    # a smaller list of the same lattices fits
    want = lat.measure([5, 6])
    rc, obs, qr, qw = noalloc(lat, [5, 6])
    assert rc == 0 and np.array_equal(obs, want[0]) and np.array_equal(qr, want[1])
