"""Grain sizes of a distributed run (the drop-in driver's draw,
nanomagmc_amd/host/grain_sizes.h): sizes ~ LogNormal(AtoLn(MEANSIZE,
SIZEDEV)) (Old/lib/main.cpp:35-38,116; functions.cpp:56-62;
stdrand.hpp:143).  The header is compiled here with g++ and its draws are
tested against the closed-form AtoLn parameters and by a Kolmogorov-Smirnov
test on 512 draws.  CPU only."""
import math
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT


@pytest.fixture(scope="module")
def draw(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("gs") / "grain_sizes_check")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "grain_sizes_check.cpp"), "-o", exe],
                   check=True)

    def run(seed, amean, asd, n, lo=0.0, hi=0.0):
        out = subprocess.run([exe, str(seed), repr(amean), repr(asd), str(n), repr(lo), repr(hi)], check=True,
                             capture_output=True, text=True).stdout.split()
        return float(out[0]), float(out[1]), np.array([float(x) for x in out[2:]])
    return run


def test_atoln_closed_form(draw):
    # MEANSIZE 10, SIZEDEV 2.5 (the grain batch's radius form): log-mean
    # ln(m^2 / sqrt(s^2 + m^2)) = 2.27227, log-sd sqrt(ln(1 + s^2/m^2)) = 0.24622
    lm, ls, _ = draw(1, 10.0, 2.5, 1)
    assert abs(lm - math.log(100.0 / math.sqrt(106.25))) < 1e-14 and abs(lm - 2.27227) < 1e-5
    assert abs(ls - math.sqrt(math.log(1.0625))) < 1e-14 and abs(ls - 0.24622) < 1e-5


@pytest.mark.parametrize("seed", [0x4E4D4300, 12345])
def test_sizes_follow_the_lognormal_law(draw, seed):
    from scipy import stats
    amean, asd, n = 10.0, 2.5, 512
    lm, ls, s = draw(seed, amean, asd, n)
    assert s.size == n and np.all(s > 0)
    ks = stats.kstest(s, stats.lognorm(ls, 0, math.exp(lm)).cdf)
    assert ks.pvalue > 1e-3, ks
    # the point of AtoLn: the sizes have mean MEANSIZE and sd SIZEDEV
    assert abs(s.mean() - amean) < 4 * asd / math.sqrt(n)
    assert abs(s.std(ddof=1) - asd) < 4 * asd / math.sqrt(2 * (n - 1)) * 1.5
    # a log-normal with the parameters swapped or unconverted is rejected
    assert stats.kstest(s, stats.lognorm(amean / 100, 0, math.exp(lm)).cdf).pvalue < 1e-6
    assert stats.kstest(s, stats.lognorm(ls, 0, amean * 1.1).cdf).pvalue < 1e-6


def test_sizes_are_seeded_and_clamped(draw):
    _, _, a = draw(7, 10.0, 2.5, 512)
    _, _, b = draw(7, 10.0, 2.5, 512)
    _, _, c = draw(8, 10.0, 2.5, 512)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    _, _, t = draw(7, 10.0, 2.5, 512, 4.0, 20.0)
    assert np.array_equal(t, np.clip(a, 4.0, 20.0))
    assert t.min() >= 4.0 and t.max() <= 20.0
