"""The benchmark batches as lattice instances of one context each (used by
bench.py and by the GPU tests that run the same batches).

* heis3d: 3-D simple-cubic classical Heisenberg, L = 64 periodic, J = 1,
  D = 0, h_z in {0, 0.05} x 32 temperatures 2.0 -> 0.8 = 64 instances.
* grains: 512 Weibull spheres (LATTSHAPE x, ISPERIO 0, WEIBULLFACT 8), radius
  scale log-normal with mean 10 and sd 2.5 (AtoLn, Old/lib/functions.cpp:
  56-62) truncated to [4, 20], in an edge-45 box, J = 1, Bloch D = 0.2
  (Old/Js/3d_nn_bloch.dat), h_z in {0, 0.02, 0.05, 0.1} x 16 temperatures
  2.0 -> 0.1 = 32768 instances, each grain cropped to its bounding box.
* ising2d: 2-D square Ising L = 64 periodic, J = 1, h = 0, 21 temperatures
  3.5 -> 1.5.

Each builder returns (lattices, present sites per instance, algorithmic
bytes per update, description).  Spins start uniformly random from the
context's Philox stream."""
import math

import numpy as np

NN3 = [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)]


def heis3d(trial="cone", replicas=1, ninst=None):
    """64 instances of the 64^3 periodic Heisenberg lattice (32 T x 2 h),
    times `replicas` independent copies of each chain (the first `ninst`
    of them if given)."""
    from .engine import Bond, Lattices
    L = 64
    temps = list(np.linspace(2.0, 0.8, 32))
    fields = [0.0, 0.05]
    bonds = [Bond(d, 1.0) for d in NN3]
    points = [(h, T) for h in fields for T in temps] * replicas
    if ninst is not None:
        points = points[:ninst]
    insts = [{"n": (L, L, L), "edge": L} for _ in points]
    lat = Lattices(3, False, True, bonds, insts, use_j=True, use_d=False, trial=trial, seed=0x4E4D4302)
    lat.set_params([1.0 / T for h, T in points], [h for h, T in points])
    lat.randomize()
    sites = [L ** 3] * lat.ninst
    desc = ("3-D cubic Heisenberg L=64 periodic, J=1, D=0, h_z in {0,0.05} x 32 T 2.0->0.8 = 64 instances%s, "
            "one checkerboard sweep per step, %s trials"
            % ("" if replicas == 1 else " x %d replicas" % replicas, trial))
    return lat, sites, 36, desc


def grain_masks(ngr=512, edge=45, wbeta=8.0, amean=10.0, asd=2.5, lo=4.0, hi=20.0, seed=0x4E4D4304):
    """Radius scales and Weibull masks (edge^3 box-scan bytes) of the grains."""
    from .engine import build_weibull
    lmean = math.log(amean * amean / math.sqrt(asd * asd + amean * amean))
    lsd = math.sqrt(math.log(1 + asd * asd / (amean * amean)))
    rng = np.random.default_rng(seed)
    sizes = np.clip(rng.lognormal(lmean, lsd, ngr), lo, hi)
    return sizes, build_weibull(3, edge, sizes, wbeta, seed=seed)


def grains(trial="sphere", replicas=1, site_range=None):
    """The 32768 grain instances; with site_range = (lo, hi) only the grains
    with lo <= present sites <= hi are kept (to time the LDS-resident and
    streamed parts of the batch apart).  The instances of grain g are
    64 g .. 64 g + 63 (4 h x 16 T) when no range is given."""
    from .engine import Bond, Lattices, crop_to_box
    edge = 45
    _, masks = grain_masks(edge=edge)
    temps = list(np.linspace(2.0, 0.1, 16))
    fields = [0.0, 0.02, 0.05, 0.1]
    insts, sites, betas, hs = [], [], [], []
    for g in range(masks.shape[0]):
        n, lo, sub = crop_to_box(masks[g], 3, edge)
        nk = int(sub.sum())
        if site_range is not None and not site_range[0] <= nk <= site_range[1]:
            continue
        for h in fields:
            for T in temps:
                insts.append({"n": tuple(n), "origin": tuple(lo), "edge": edge, "mask": sub})
                sites.append(nk)
                betas.append(1.0 / T)
                hs.append(h)
    bonds = [Bond(d, 1.0, tuple(0.2 * x for x in d)) for d in NN3]
    lat = Lattices(3, False, False, bonds, insts, use_j=True, use_d=True, trial=trial, seed=0x4E4D4304)
    lat.set_params(betas, hs)
    lat.randomize()
    desc = ("512 Weibull grains (radius log-normal 10+-2.5 in [4,20], factor 8, open) x 4 h x 16 T = "
            "32768 instances, Heisenberg J=1 + Bloch D=0.2, one sweep of every instance per step, %s trials"
            % trial)
    if site_range is not None:
        desc += "; only the %d instances of grains with %d..%d present sites" % (len(insts), site_range[0],
                                                                                 site_range[1])
    return lat, sites, 36, desc


def ising2d(trial="sphere", replicas=1):
    from .engine import Bond, Lattices
    L = 64
    temps = list(np.linspace(3.5, 1.5, 21))
    bonds = [Bond((0, 1), 1.0), Bond((0, -1), 1.0), Bond((1, 0), 1.0), Bond((-1, 0), 1.0)]
    lat = Lattices(2, True, True, bonds, [{"n": (L, L), "edge": L}] * len(temps), seed=0x4E4D4301)
    lat.set_params([1.0 / T for T in temps], 0.0)
    lat.randomize()
    desc = "2-D square Ising L=64 periodic, J=1, h=0, 21 T 3.5->1.5, one checkerboard sweep per step"
    return lat, [L * L] * lat.ninst, 3, desc


WORKLOADS = {"heis3d": heis3d, "grains": grains, "ising2d": ising2d}
