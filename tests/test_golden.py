"""The committed golden vectors (tests/golden/vectors.npz, written by
tests/golden/make_golden.py from the oracle) against the oracle, bit for bit,
and against closed forms that do not use the oracle:

* Ising dE (thermodynamics.cpp:10-69 with integer spins): 2 s (h + sum_j J_j s_j)
  for a flip, 0 for the trial equal to the spin;
* Heisenberg dE: (H + sum_j J_j s_j + s_j x D_j) . (s - t) with neighbours
  found here by the box offsets (wrapped if periodic);
* the replayed decisions: accepted iff dE <= 0 or -dE beta > log(k 2^-32),
  with dE and the spins advanced here step by step;
* the skyrmion charge: -1 within 0.01 in every slice (heis_test.hpp:378-393).
"""
import math
import os

import numpy as np
import pytest

from tests.golden import make_golden


@pytest.fixture(scope="module")
def gold():
    with np.load(make_golden.OUT, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_the_golden_vectors(orc, gold):
    v = make_golden.vectors(orc)
    assert sorted(v) == sorted(gold)
    for k in v:
        assert np.array_equal(np.asarray(v[k]), gold[k]), k


def _neighbours(dim, edge, periodic, dirs):
    """Site index of each bond's neighbour (-1: outside an open box), sites in
    box-scan order with the last coordinate fastest."""
    g = np.indices((edge,) * dim).reshape(dim, -1).T
    out = np.full((len(g), len(dirs)), -1)
    for k, d in enumerate(dirs):
        q = g + np.asarray(d[:dim])[None, :]
        if periodic:
            q %= edge
            ok = np.ones(len(g), dtype=bool)
        else:
            ok = np.all((q >= 0) & (q < edge), axis=1)
        idx = np.ravel_multi_index(tuple(np.clip(q, 0, edge - 1).T), (edge,) * dim)
        out[ok, k] = idx[ok]
    return out


@pytest.mark.parametrize("name", sorted(make_golden.DE_CASES))
def test_delta_e_vectors_against_the_closed_form(gold, name):
    dim, edge, periodic, ising, use_d, _ = make_golden.DE_CASES[name]
    p = "de_%s_" % name
    s, sites, trials, want = gold[p + "spins"], gold[p + "sites"], gold[p + "trials"], gold[p + "want"]
    dirs, J, D, h = gold[p + "dirs"], gold[p + "J"], gold[p + "D"], float(gold[p + "h"])
    nb = _neighbours(dim, edge, periodic, dirs)
    for i, j in enumerate(sites):
        V = np.zeros(3)
        for k in range(len(J)):
            if nb[j, k] < 0:
                continue
            n = s[nb[j, k]]
            V += J[k] * n + (np.cross(n, D[k]) if use_d else 0.0)
        if ising:
            H = np.array([h, 0.0, 0.0])
            exact = 0.0 if trials[i, 0] == s[j, 0] else 2.0 * s[j, 0] * (h + V[0])
            assert want[i] == exact, (i, want[i], exact)
        else:
            H = np.array([0.0, 0.0, h])
            exact = float(np.dot(H + V, s[j] - trials[i]))
            assert abs(want[i] - exact) <= 1e-12 * (1.0 + abs(exact)), (i, want[i], exact)
            assert gold[p + "scale"][i] == pytest.approx(np.linalg.norm(H + V) * np.linalg.norm(s[j] - trials[i]),
                                                         rel=1e-12)


def test_replay_vector_against_a_step_by_step_update(gold):
    p = "rp_ising2d_"
    s = gold[p + "spins"][:, 0].copy()
    T, h = float(gold[p + "T"]), float(gold[p + "h"])
    edge = 16
    nb = _neighbours(2, edge, True, make_golden.nn2d()[0])
    acc = []
    for i, t, k in zip(gold[p + "sites"], gold[p + "trials"], gold[p + "k"]):
        dE = 0.0 if t == s[i] else 2.0 * s[i] * (h + s[nb[i]].sum())
        u = int(k) / 4294967296.0
        a = dE <= 0 or (-dE / T) > (math.log(u) if u > 0 else -math.inf)
        if a:
            s[i] = t
        acc.append(1 if a else 0)
    assert np.array_equal(np.array(acc, dtype=np.uint8), gold[p + "accept"])
    assert 0 < gold[p + "accept"].sum() < len(acc)


def test_skyrmion_charge_vector(gold):
    q = gold["tc_skyrm30"]
    assert q.shape == (30,)
    assert np.all(np.abs(q + 1) < 0.01)


def test_golden_file_is_small():
    assert os.path.getsize(make_golden.OUT) < 100 * 1024
