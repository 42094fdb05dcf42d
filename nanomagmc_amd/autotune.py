"""Launch-shape autotuner for nmc_sweep.

The engine has two ways to sweep an instance (LDS-resident whole sweeps, or
streaming colour passes) and a tunable workgroup size for the LDS-resident
kernel.  Results do not depend on either (both draw the same random numbers),
so the tuner is free to pick whatever is fastest for a batch: it times a few
sweeps of every candidate on the batch itself and keeps the fastest.  The
decision is cached in a JSON file keyed by a signature of the batch.
The timing sweeps are real sweeps of the chains (use them as equilibration).
"""
import json
import os
import time

CANDIDATES = [("auto", 0), ("auto", 256), ("auto", 512), ("auto", 1024), ("stream", 0)]
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "nanomagmc_amd", "autotune.json")


def _time(lat, path, block, sweeps):
    lat.set_path(path)
    lat.set_lds_block(block)
    lat.equil(1)                 # builds the work lists outside the timed region
    lat.sync()
    t0 = time.perf_counter()
    lat.equil(sweeps)
    lat.sync()
    return time.perf_counter() - t0


def tune(lat, signature, sweeps=4, cache_path=DEFAULT_CACHE, candidates=CANDIDATES, force=False):
    """Pick (path, lds_block) for `lat`; returns (choice, timings)."""
    cache = {}
    if os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except (OSError, ValueError):
            cache = {}
    if not force and signature in cache:
        path, block = cache[signature]["choice"]
        lat.set_path(path)
        lat.set_lds_block(block)
        return (path, block), cache[signature].get("timings", {})
    timings = {}
    for path, block in candidates:
        timings["%s/%d" % (path, block)] = _time(lat, path, block, sweeps)
    best = min(candidates, key=lambda c: timings["%s/%d" % c])
    lat.set_path(best[0])
    lat.set_lds_block(best[1])
    cache[signature] = {"choice": list(best), "timings": timings}
    try:
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        json.dump(cache, open(cache_path, "w"), indent=1)
    except OSError:
        pass
    return best, timings
