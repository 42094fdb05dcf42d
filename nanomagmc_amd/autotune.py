"""Launch-shape autotuner for nmc_sweep.

The engine sweeps an instance either LDS-resident (whole sweeps in one
workgroup: the dense box or the compacted present sites) or by streaming
colour passes, and three launch shapes are free: the path (AUTO, streaming
only, or open shaped lattices streamed from compacted copies), the workgroup size of the LDS-resident kernels and the tile
size of the streaming passes over colour-split boxes (64-work-item quarter
tiles or 256-work-item full tiles).  Results do not depend on any of them
(every path draws the same random numbers per site), so the tuner is free to
pick whatever is fastest for a batch: it times a few sweeps of every
candidate on the batch itself and keeps the fastest.  The decision is cached
in a JSON file keyed by a signature of the batch.  The timing sweeps are
real sweeps of the chains (use them as equilibration).
"""
import json
import os
import time

# (path, LDS workgroup size, streaming tile work-items); 0 = the engine's choice
CANDIDATES = [("auto", 0, 0), ("auto", 256, 0), ("auto", 512, 0), ("auto", 1024, 0), ("auto", 0, 64),
              ("auto", 0, 256), ("stream", 0, 0), ("stream", 0, 64), ("stream", 0, 256), ("cstream", 0, 0)]
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "nanomagmc_amd", "autotune.json")


def apply(lat, choice):
    path, block, sblock = choice
    lat.set_path(path)
    lat.set_lds_block(block)
    lat.set_stream_block(sblock)


def key(choice):
    return "%s/%d/%d" % tuple(choice)


def _time(lat, choice, sweeps, reps=2):
    apply(lat, choice)
    lat.equil(1)                 # builds the work lists (and warms the kernels) outside the timed region
    lat.sync()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        lat.equil(sweeps)
        lat.sync()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(lat, signature, sweeps=8, cache_path=DEFAULT_CACHE, candidates=CANDIDATES, force=False):
    """Pick (path, lds_block, stream_block) for `lat` and apply it; returns
    (choice, seconds per candidate for `sweeps` sweeps)."""
    cache = {}
    if os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except (OSError, ValueError):
            cache = {}
    if not force and signature in cache and len(cache[signature].get("choice", ())) == 3:
        choice = tuple(cache[signature]["choice"])
        apply(lat, choice)
        return choice, cache[signature].get("timings", {})
    timings = {}
    for c in candidates:
        timings[key(c)] = _time(lat, c, sweeps)
    best = min(candidates, key=lambda c: timings[key(c)])
    apply(lat, best)
    cache[signature] = {"choice": list(best), "timings": timings}
    try:
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        json.dump(cache, open(cache_path, "w"), indent=1)
    except OSError:
        pass
    return best, timings
