"""Helpers shared by the GPU parity tests (construct matching GPU contexts
and oracle fields from one description)."""
import numpy as np

from tests.fixtures import box_locs


def engine_bonds(d, J, D):
    from nanomagmc_amd import Bond
    return [Bond(d[k], J[k], D[k]) for k in range(len(J))]


def mask_from_locs(locs, dim, n_box, origin=(0, 0, 0)):
    m = np.zeros(n_box[:dim], dtype=np.uint8)
    rel = locs[:, :dim] - np.asarray(origin[:dim])[None, :]
    m[tuple(rel.T)] = 1
    return m.reshape(-1)


def make_pair(orc, dim, edge, periodic, bd, J, D, ising, use_d=False, locs=None, trial="sphere",
              seed=0x4E4D4301, ninst=1, n_box=None, origin=(0, 0, 0)):
    """A GPU context with `ninst` copies of one lattice and the oracle Field."""
    from nanomagmc_amd import Lattices
    full = locs is None
    if full:
        locs = box_locs(dim, edge)
    if n_box is None:
        n_box = (edge,) * dim
    inst = {"n": tuple(n_box), "origin": tuple(origin), "edge": edge}
    if not full:
        inst["mask"] = mask_from_locs(locs, dim, n_box, origin)
    lat = Lattices(dim, ising, periodic, engine_bonds(bd, J, D), [dict(inst) for _ in range(ninst)],
                   use_j=True, use_d=use_d, trial=trial, seed=seed)
    f = orc.Field(locs, dim, edge, periodic, bd, J, D)
    return lat, f
