"""End-to-end runs of the C++ host driver nmc_run on tiny inputs: INPUT.dat
keys, protocol path order, HDF5 layout (Old/lib/output.cpp:80-458) and the
/Complete resume flags.  Needs the GPU and the HDF5 command-line tools."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(ROOT, "nanomagmc_amd", "nmc_run")
H5DUMP = "/opt/conda/bin/h5dump"
H5LS = "/opt/conda/bin/h5ls"


def run(case_dir, timeout=120):
    r = subprocess.run([RUNNER, "INPUT.dat"], cwd=case_dir, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def h5_shapes(path):
    out = subprocess.run([H5LS, "-r", path], capture_output=True, text=True, check=True).stdout
    shapes = {}
    for line in out.splitlines():
        m = re.match(r"^(\S+)\s+Dataset \{([^}]*)\}", line)
        if m:
            shapes[m.group(1)] = tuple(int(x.split("/")[0]) for x in m.group(2).split(","))
    return shapes


def h5_read(path, name, dtype=None):
    """Raw little-endian dump of one dataset (h5dump -b LE)."""
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".bin") as tmp:
        subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", tmp.name, path], capture_output=True, text=True,
                       check=True)
        raw = open(tmp.name, "rb").read()
    if dtype is None:
        dtype = np.int32 if name.endswith("Complete") else np.float32
    return np.frombuffer(raw, dtype=dtype).astype(np.float64)


def make_case(tmp_path, **keys):
    from nanomagmc_amd.inputs import write_case
    base = dict(ISISING=1, EXCHANGE=1, DMISTREN=0, ISDISTRIB=0, SIZE=8, LATTSHAPE="s", NLATTS=2, BOLTZMANN=1,
                ISPERIO=1, SAMPSWEEPS=2, EQSWEEPS=10, NSAMPS=5, PROTOCOL=1, INTERACTIONS="2d_nn.dat")
    temps = keys.pop("temps", [3.0, 2.0, 1.5])
    fields = keys.pop("fields", [0.0, 0.1])
    base.update(keys)
    write_case(str(tmp_path), base, temps, fields)
    return str(tmp_path)


@pytest.mark.parametrize("protocol", [1, 2, 4])
def test_ising_layout(tmp_path, protocol):
    d = make_case(tmp_path, PROTOCOL=protocol)
    r = run(d)
    assert "3 temperatures" in r.stdout and "2 field strengths" in r.stdout
    path = os.path.join(d, "Output", "J_1-s_8-k_1-K_0-sh_s-per_1-pr_%d.h5" % protocol)
    assert os.path.exists(path)
    sh = h5_shapes(path)
    for n in ("mags", "energies", "sub_mags", "sub4_mags"):
        assert sh["/Fulldat/" + n] == (2, 3, 10)
    assert "/Fulldat/mag_xs" not in sh and "/Fulldat/top_chars" not in sh
    assert sh["/Fulldat/nums"] == (5,)
    assert sh["/Ts"] == (3,) and sh["/Hs"] == (2,)
    v1 = 2 if protocol == 1 else 3
    assert sh["/Complete"] == (v1, 2)
    assert sh["/Av_Latt/T_0-H_0"] == (8, 8, 1) and sh["/Sing_Latt/T_2-H_1"] == (8, 8, 1)
    E = h5_read(path, "/Fulldat/energies").reshape(2, 3, 10)
    Hs = [0.0, 0.1]
    # H index 0 is h = 0: E = -sum s_i s_j over 128 bonds, a multiple of 4
    assert np.all(np.abs(E[0]) <= 2 * 64) and np.all(E[0] % 4 == 0)
    assert np.all(np.abs(E[1]) <= 2 * 64 + Hs[1] * 64 + 1e-4)
    M = h5_read(path, "/Fulldat/mags").reshape(2, 3, 10)
    assert np.all(M <= 64) and np.all(M >= 0)
    # every (H, T, lattice) slot was written: the ordered phase at T = 1.5
    # has |M| close to N
    assert M[0, 2].mean() > 40
    comp = h5_read(path, "/Complete").reshape(v1, 2)
    if protocol != 4:
        assert np.all(comp == 1)
    Ts = h5_read(path, "/Ts")
    assert np.allclose(Ts, [3.0, 2.0, 1.5])
    sing = h5_read(path, "/Sing_Latt/T_2-H_0")
    assert set(np.unique(sing)) <= {-1.0, 1.0}


def test_resume_skips_completed_chains(tmp_path):
    d = make_case(tmp_path, PROTOCOL=2)
    run(d)
    path = os.path.join(d, "Output", "J_1-s_8-k_1-K_0-sh_s-per_1-pr_2.h5")
    before = h5_read(path, "/Fulldat/energies")
    r = run(d)   # everything complete: the v2 walks are skipped, data unchanged
    assert "Printing" not in r.stdout
    assert np.array_equal(before, h5_read(path, "/Fulldat/energies"))


def test_heisenberg_charges_and_components(tmp_path):
    d = make_case(tmp_path, ISISING=0, DMISTREN=0.3, INTERACTIONS="2d_nn_neel.dat", SIZE=12, NLATTS=1,
                  temps=[1.0, 0.5], fields=[0.0])
    run(d)
    path = os.path.join(d, "Output", "J_1-s_12-k_1-K_0.3-sh_s-per_1-pr_1.h5")
    sh = h5_shapes(path)
    assert sh["/Fulldat/top_chars"] == (1, 2, 12, 5)
    assert sh["/Fulldat/mag_zs"] == (1, 2, 5)
    assert sh["/Av_Latt/T_1-H_0"] == (12, 12, 3)
    mx, my, mz = (h5_read(path, "/Fulldat/mag_%ss" % c) for c in "xyz")
    m = h5_read(path, "/Fulldat/mags")
    assert np.allclose(np.sqrt(mx ** 2 + my ** 2 + mz ** 2), m, rtol=1e-5, atol=1e-4)
    qw = h5_read(path, "/Fulldat/top_chars_wrap").reshape(1, 2, 12, 5)
    assert np.allclose(qw[:, :, 0, :], np.round(qw[:, :, 0, :]), atol=1e-3)   # integer on the torus
    assert np.all(qw[:, :, 1:, :] == 0)


def test_weibull_grains_distributed(tmp_path):
    d = make_case(tmp_path, ISISING=0, DMISTREN=0.2, INTERACTIONS="3d_nn_bloch.dat", LATTSHAPE="x",
                  ISDISTRIB=1, MEANSIZE=3, SIZEDEV=0.5, WEIBULLFACT=8, ISPERIO=0, NLATTS=3,
                  temps=[1.0], fields=[0.0, 0.05])
    run(d)
    path = os.path.join(d, "Output", "J_1-am_3-asd_0.5-k_1-K_0.2-sh_x-per_0-pr_1.h5")
    sh = h5_shapes(path)
    assert sh["/Fulldat/energies"] == (2, 1, 15)
    assert "/Fulldat/top_chars" not in sh and "/Av_Latt/T_0-H_0" not in sh
    m = h5_read(path, "/Fulldat/mags")
    assert np.all(m > 0)


def test_series_blocks_sit_at_their_lattice(tmp_path):
    # a step's chains are written one hyperslab per run of lattices: lattice
    # 0's block of the component series must sum to its average lattice (the
    # plain sum of its sampled spins, state::add_to_av), and the other
    # lattices' blocks are other chains
    d = make_case(tmp_path, ISISING=0, DMISTREN=0.3, INTERACTIONS="2d_nn_neel.dat", SIZE=10, NLATTS=3,
                  temps=[1.5, 0.7], fields=[0.0, 0.1])
    run(d)
    path = os.path.join(d, "Output", "J_1-s_10-k_1-K_0.3-sh_s-per_1-pr_1.h5")
    ns, nl = 5, 3
    comps = [h5_read(path, "/Fulldat/mag_%ss" % c).reshape(2, 2, nl * ns) for c in "xyz"]
    for H in range(2):
        for T in range(2):
            av = h5_read(path, "/Av_Latt/T_%d-H_%d" % (T, H)).reshape(-1, 3).sum(axis=0)
            blk = np.array([c[H, T, :ns].sum() for c in comps])
            assert np.allclose(blk, av, rtol=1e-4, atol=1e-2), (H, T, blk, av)
            for k in (1, 2):
                other = np.array([c[H, T, k * ns:(k + 1) * ns] for c in comps])
                assert not np.array_equal(other, np.array([c[H, T, :ns] for c in comps]))


def _h5_write_block(path, name, start, count, values, integer=False):
    """Overwrite one hyperslab of a dataset through the HDF5 C library."""
    import ctypes as C
    h5 = C.CDLL("/opt/conda/lib/libhdf5.so")
    h5.H5open()
    hid = C.c_int64
    for f in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Screate_simple"):
        getattr(h5, f).restype = hid
    ftype = hid.in_dll(h5, "H5T_NATIVE_INT_g" if integer else "H5T_NATIVE_FLOAT_g")
    n = len(start)
    st = (C.c_uint64 * n)(*start)
    ct = (C.c_uint64 * n)(*count)
    buf = np.ascontiguousarray(values, dtype=np.int32 if integer else np.float32)
    f = hid(h5.H5Fopen(path.encode(), C.c_uint(1), hid(0)))
    ds = hid(h5.H5Dopen2(f, name.encode(), hid(0)))
    sp = hid(h5.H5Dget_space(ds))
    assert h5.H5Sselect_hyperslab(sp, 0, st, None, ct, None) >= 0
    mem = hid(h5.H5Screate_simple(n, ct, None))
    assert h5.H5Dwrite(ds, ftype, mem, sp, hid(0), buf.ctypes.data_as(C.c_void_p)) >= 0
    for h, fn in ((mem, "H5Sclose"), (sp, "H5Sclose"), (ds, "H5Dclose"), (f, "H5Fclose")):
        getattr(h5, fn)(h)


def test_resume_rewrites_only_incomplete_chains(tmp_path):
    # two chains marked incomplete and their series blanked: a rerun must
    # rewrite exactly those blocks (runs of lattices split around the
    # complete ones) and reproduce the first run's values bit for bit
    d = make_case(tmp_path, PROTOCOL=2, NLATTS=3)
    run(d)
    path = os.path.join(d, "Output", "J_1-s_8-k_1-K_0-sh_s-per_1-pr_2.h5")
    names = ["/Fulldat/energies", "/Fulldat/mags", "/Fulldat/sub_mags", "/Fulldat/sub4_mags",
             "/Fulldat/energies_s"]
    before = {n: h5_read(path, n) for n in names + ["/Complete"]}
    ns, nl = 5, 3
    sh = h5_shapes(path)
    nH, nT = sh["/Fulldat/energies"][:2]
    # PROTOCOL 2: v1 = T index, v2 = H index; chain (v1 = T, lattice k)
    for T, k in ((1, 1), (0, 2)):
        _h5_write_block(path, "/Complete", (T, k), (1, 1), [0], integer=True)
        for n in names:
            for H in range(nH):
                _h5_write_block(path, n, (H, T, k * ns), (1, 1, ns), np.zeros(ns))
    blank = h5_read(path, "/Fulldat/energies").reshape(nH, nT, nl * ns)
    assert np.all(blank[:, 1, ns:2 * ns] == 0) and np.all(blank[:, 0, 2 * ns:] == 0)
    r = run(d)
    assert "Printing" in r.stdout
    for n in names + ["/Complete"]:
        assert np.array_equal(h5_read(path, n), before[n]), n
