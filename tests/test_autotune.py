"""Launch-shape autotuner (nanomagmc_amd.autotune) on a stand-in batch whose
sweep time depends on the chosen path, workgroup size and tile size.  CPU
only (tests/test_gpu_autotune.py runs it on a real batch)."""
import json
import time

from nanomagmc_amd import autotune


class FakeBatch:
    """Records the configuration; equil() sleeps for that configuration's cost."""

    def __init__(self, cost):
        self.cost = cost
        self.path, self.block, self.sblock = "auto", 0, 0
        self.sweeps = 0

    def set_path(self, path):
        self.path = path

    def set_lds_block(self, block):
        self.block = block

    def set_stream_block(self, block):
        self.sblock = block

    def config(self):
        return (self.path, self.block, self.sblock)

    def equil(self, n):
        self.sweeps += n
        time.sleep(self.cost[self.config()] * n)

    def sync(self):
        pass


def costs(best):
    c = {cand: 2e-3 for cand in autotune.CANDIDATES}
    c[best] = 2e-4
    return c


def test_picks_the_fastest_candidate_and_caches_it(tmp_path):
    cache = tmp_path / "tune.json"
    lat = FakeBatch(costs(("auto", 512, 0)))
    choice, timings = autotune.tune(lat, "sig-a", sweeps=3, cache_path=str(cache))
    assert tuple(choice) == ("auto", 512, 0) and lat.config() == ("auto", 512, 0)
    assert set(timings) == {autotune.key(c) for c in autotune.CANDIDATES}
    assert lat.sweeps == len(autotune.CANDIDATES) * (1 + 2 * 3)   # 1 untimed + 2 x 3 timed sweeps each
    saved = json.loads(cache.read_text())
    assert saved["sig-a"]["choice"] == ["auto", 512, 0]
    # a cached signature is applied without timing anything
    lat2 = FakeBatch(costs(("stream", 0, 64)))
    choice2, _ = autotune.tune(lat2, "sig-a", cache_path=str(cache))
    assert tuple(choice2) == ("auto", 512, 0) and lat2.sweeps == 0 and lat2.config() == ("auto", 512, 0)
    # force re-times and replaces the entry
    choice3, _ = autotune.tune(lat2, "sig-a", sweeps=2, cache_path=str(cache), force=True)
    assert tuple(choice3) == ("stream", 0, 64) and lat2.config() == ("stream", 0, 64)
    assert json.loads(cache.read_text())["sig-a"]["choice"] == ["stream", 0, 64]


def test_unreadable_or_old_cache_is_ignored(tmp_path):
    cache = tmp_path / "tune.json"
    cache.write_text("{not json")
    lat = FakeBatch(costs(("auto", 0, 256)))
    choice, _ = autotune.tune(lat, "sig-b", sweeps=2, cache_path=str(cache))
    assert tuple(choice) == ("auto", 0, 256)
    assert json.loads(cache.read_text())["sig-b"]["choice"] == ["auto", 0, 256]
    # an entry of two fields (path, block) is re-tuned
    cache.write_text(json.dumps({"sig-c": {"choice": ["auto", 512], "timings": {}}}))
    choice, _ = autotune.tune(lat, "sig-c", sweeps=1, cache_path=str(cache))
    assert tuple(choice) == ("auto", 0, 256)
