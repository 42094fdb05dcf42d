"""The autotuner on real batches (needs an MI355X): every candidate launch
shape gives the same lattices bit for bit (the random numbers of a site do
not depend on the launch), and the tuner's choice is the candidate that was
measured fastest.  Batches: shaped open grains (compacted LDS-resident and
span-streamed) and a set of periodic 3-D lattices (LDS-resident and streamed
over the box)."""
import numpy as np
import pytest

from tests.gpu_util import engine_bonds

pytestmark = pytest.mark.gpu


def grain_batch(orc, bloch3d, seed):
    from nanomagmc_amd import Lattices, build_weibull, crop_to_box
    bd, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.2)
    masks = build_weibull(3, 41, [5.0, 9.0, 18.0], 8.0, seed=0x4E4D4340)
    insts = []
    for m in masks:
        n, lo, sub = crop_to_box(m, 3, 41)
        insts += [{"n": tuple(n), "origin": tuple(lo), "edge": 41, "mask": sub}] * 4
    lat = Lattices(3, False, False, engine_bonds(bd, J, D), insts, use_d=True, seed=seed)
    lat.set_params(np.linspace(0.5, 5.0, len(insts)), 0.05)
    lat.randomize()
    return lat


def periodic_batch(seed):
    from nanomagmc_amd import Lattices
    d3 = [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)]
    insts = [{"n": (8, 8, 8), "edge": 8}] * 3 + [{"n": (32, 32, 32), "edge": 32}] * 2
    lat = Lattices(3, False, True, engine_bonds(d3, [1.0] * 6, [(0.0, 0.0, 0.0)] * 6), insts, trial="cone", seed=seed)
    lat.set_params([0.5, 1.0, 2.0, 0.7, 1.4], 0.05)
    lat.set_cone([0.8] * 5)
    lat.randomize()
    return lat


@pytest.mark.parametrize("kind", ["grains", "periodic"])
def test_every_candidate_gives_the_same_lattices(orc, bloch3d, kind):
    from nanomagmc_amd import autotune
    ref = None
    for c in autotune.CANDIDATES:
        lat = grain_batch(orc, bloch3d, 7) if kind == "grains" else periodic_batch(7)
        autotune.apply(lat, c)
        lat.equil(9)
        spins = [lat.get_spins(i) for i in range(lat.ninst)]
        acc = lat.accepts()
        if ref is None:
            ref = (spins, acc)
            continue
        for i in range(lat.ninst):
            assert np.array_equal(spins[i], ref[0][i]), (c, i)
        assert np.array_equal(acc, ref[1]), c


def test_choice_is_the_measured_fastest(orc, bloch3d, tmp_path):
    from nanomagmc_amd import autotune
    lat = grain_batch(orc, bloch3d, 11)
    choice, timings = autotune.tune(lat, "gpu-test", sweeps=6, cache_path=str(tmp_path / "t.json"), force=True)
    assert set(timings) == {autotune.key(c) for c in autotune.CANDIDATES}
    assert timings[autotune.key(choice)] == min(timings.values())
    # the chosen shape stays in force: the path it names decides the kernels
    kinds = set(lat.sweep_kinds().tolist())
    from nanomagmc_amd._lib import NMC_KIND_COMPACT_STREAM, NMC_KIND_LDS_COMPACT, NMC_KIND_SPANS
    want = {"stream": {NMC_KIND_SPANS}, "cstream": {NMC_KIND_COMPACT_STREAM},
            "auto": {NMC_KIND_LDS_COMPACT, NMC_KIND_COMPACT_STREAM}}
    assert kinds == want[choice[0]]
