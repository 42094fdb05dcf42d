"""Throughput of the Metropolis hot path on one MI355X.

Workload (the lattice the headline metric is quoted on): 3-D simple-cubic
classical Heisenberg, L = 64 periodic, J = 1, D = 0, fp32 unit-vector spins
drawn uniformly on the sphere, h_z in {0, 0.05} x 32 temperatures 2.0 -> 0.8:
64 lattice instances resident in HBM, swept together.  One step = one
checkerboard sweep of every instance (64 * 64^3 = 16.8M attempted single-site
updates) with cone-limited trial moves whose half-angles were adapted toward
50% acceptance during warm-up and then frozen.

Prints one JSON line (see the README of the measurement contract in this
repository's documentation).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

L = 64
TEMPS = list(np.linspace(2.0, 0.8, 32))
FIELDS = [0.0, 0.05]
BYTES_PER_UPDATE = 36          # fp32 x 3: read as target, read as neighbour, written once
HBM_PEAK = 8.0e12


def make_lattices(trial="cone"):
    from nanomagmc_amd import Bond, Lattices
    bonds = [Bond((0, 0, 1), 1.0), Bond((0, 0, -1), 1.0), Bond((0, 1, 0), 1.0),
             Bond((0, -1, 0), 1.0), Bond((1, 0, 0), 1.0), Bond((-1, 0, 0), 1.0)]
    insts = [{"n": (L, L, L), "edge": L} for _ in range(len(TEMPS) * len(FIELDS))]
    lat = Lattices(3, False, True, bonds, insts, use_j=True, use_d=False, trial=trial, seed=0x4E4D4302)
    betas, hs = [], []
    for h in FIELDS:
        for T in TEMPS:
            betas.append(1.0 / T)
            hs.append(h)
    lat.set_params(betas, hs)
    lat.randomize()
    return lat


def adapt_cone(lat, theta, sweeps):
    """One adaptation round: sweep, read acceptance, scale half-angles."""
    lat.reset_accepts()
    lat.equil(sweeps)
    acc = lat.accepts().astype(np.float64) / (sweeps * L ** 3)
    theta = np.clip(theta * np.clip(acc / 0.5, 0.5, 2.0), 0.02, math.pi)
    lat.set_cone(theta.astype(np.float32))
    return theta, acc


def cpu_baseline(budget_s=12.0):
    """The reference algorithm (oracle port) on this host: random-site
    single-site Metropolis over AoS double spins and vector<vector<int>>
    neighbour lists, mt19937, uniform-sphere trials (Old/lib/state.cpp:171-203)
    on the same 64^3 periodic lattice at T = 1.4, h = 0."""
    from oracle import oracle as orc
    orc.build()
    from tests.fixtures import box_locs
    d, J, D = orc.nn_bonds(3)
    f = orc.Field(box_locs(3, L), 3, L, 1, d, J, D)
    threads = max(1, min(16, os.cpu_count() or 1))
    probe, _ = f.bench_reference(False, 0.0, 1 / 1.4, 1, 1_000_000)
    attempts = int(max(1e6, probe * budget_s / 1.3))
    rate, secs = f.bench_reference(False, 0.0, 1 / 1.4, threads, attempts)
    return {"value": rate, "unit": "spin-updates/s", "cores": threads, "kind": "port",
            "sample": "%d threads x %d attempts (one chain per thread) on the 64^3 periodic Heisenberg "
                      "lattice, T=1.4, h=0, uniform-sphere trials; %.1f s" % (threads, attempts, secs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--trial", default="cone", choices=["cone", "sphere"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs independent chains per GPU and this "
              "build does not launch multi-GPU runs (use --gpus 1).", file=sys.stderr)
        sys.exit(2)
    import torch
    lat = make_lattices(args.trial)
    sites = lat.ninst * L ** 3
    # warm-up: adapt cone half-angles toward 50% acceptance, then freeze them
    theta = np.full(lat.ninst, 1.0)
    if args.trial == "cone":
        lat.set_cone(theta.astype(np.float32))
    done = 0
    while done < args.warmup:
        k = min(10, args.warmup - done)
        if args.trial == "cone":
            theta, acc = adapt_cone(lat, theta, k)
        else:
            lat.equil(k)
        done += k
    lat.reset_accepts()
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(lat.stream())
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    lat.equil(args.steps)              # 2 * steps colour-pass launches + 1 counter bump
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    acc = lat.accepts().astype(np.float64).sum() / (args.steps * sites)
    updates = args.steps * sites
    value = updates / wall
    launch_s = gpu_s / (2 * args.steps)
    achieved = (sites / 2) * BYTES_PER_UPDATE / launch_s
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_heis3d_L64x64.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
    out = {
        "metric": "Metropolis spin-updates/s, 1 GPU",
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (spins uniform on S^2 from Philox, seed 0x4E4D4302)",
        "config": {"workload": "3-D cubic Heisenberg L=64 periodic, J=1, D=0, h_z in {0,0.05} x 32 T "
                               "2.0->0.8 = 64 instances, one checkerboard sweep per step, %s trials"
                               % args.trial,
                   "instances": lat.ninst, "sites": sites, "acceptance": acc},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "k_sweep_heis (one colour pass: %d updates, %d B/update algorithmic)"
                               % (sites // 2, BYTES_PER_UPDATE),
                     "avg_launch_us": launch_s * 1e6},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
