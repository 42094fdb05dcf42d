"""Launch-shape autotuner (nanomagmc_amd.autotune) on a stand-in batch whose
sweep time depends on the chosen path and workgroup size.  CPU only."""
import json
import time

from nanomagmc_amd import autotune


class FakeBatch:
    """Records the configuration; equil() sleeps for that configuration's cost."""

    def __init__(self, cost):
        self.cost = cost
        self.path, self.block = "auto", 0
        self.sweeps = 0

    def set_path(self, path):
        self.path = path

    def set_lds_block(self, block):
        self.block = block

    def equil(self, n):
        self.sweeps += n
        time.sleep(self.cost[(self.path, self.block)] * n)

    def sync(self):
        pass


def costs(best):
    c = {cand: 2e-3 for cand in autotune.CANDIDATES}
    c[best] = 2e-4
    return c


def test_picks_the_fastest_candidate_and_caches_it(tmp_path):
    cache = tmp_path / "tune.json"
    lat = FakeBatch(costs(("auto", 512)))
    choice, timings = autotune.tune(lat, "sig-a", sweeps=3, cache_path=str(cache))
    assert tuple(choice) == ("auto", 512) and (lat.path, lat.block) == ("auto", 512)
    assert set(timings) == {"%s/%d" % c for c in autotune.CANDIDATES}
    assert lat.sweeps == len(autotune.CANDIDATES) * 4          # 1 untimed + 3 timed sweeps each
    saved = json.loads(cache.read_text())
    assert saved["sig-a"]["choice"] == ["auto", 512]
    # a cached signature is applied without timing anything
    lat2 = FakeBatch(costs(("stream", 0)))
    choice2, _ = autotune.tune(lat2, "sig-a", cache_path=str(cache))
    assert tuple(choice2) == ("auto", 512) and lat2.sweeps == 0 and (lat2.path, lat2.block) == ("auto", 512)
    # force re-times and replaces the entry
    choice3, _ = autotune.tune(lat2, "sig-a", sweeps=2, cache_path=str(cache), force=True)
    assert tuple(choice3) == ("stream", 0) and (lat2.path, lat2.block) == ("stream", 0)
    assert json.loads(cache.read_text())["sig-a"]["choice"] == ["stream", 0]


def test_unreadable_cache_is_ignored(tmp_path):
    cache = tmp_path / "tune.json"
    cache.write_text("{not json")
    lat = FakeBatch(costs(("auto", 1024)))
    choice, _ = autotune.tune(lat, "sig-b", sweeps=2, cache_path=str(cache))
    assert tuple(choice) == ("auto", 1024)
    assert json.loads(cache.read_text())["sig-b"]["choice"] == ["auto", 1024]
