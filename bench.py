"""Throughput of the Metropolis hot path on one MI355X.

Default workload (the lattice the headline metric is quoted on): 3-D
simple-cubic classical Heisenberg, L = 64 periodic, J = 1, D = 0, fp32
unit-vector spins drawn uniformly on the sphere, h_z in {0, 0.05} x 32
temperatures 2.0 -> 0.8: 64 lattice instances resident in HBM, swept
together.  One step = one checkerboard sweep of every instance
(64 * 64^3 = 16.8M attempted single-site updates) with cone-limited trial
moves whose half-angles are adapted toward 50% acceptance during warm-up and
then frozen.

Other workloads (--workload): `grains` (512 Weibull spheres x 64 (T, h)
points = 32768 uneven instances, Heisenberg + Bloch DMI, open boundaries),
`ising2d` (2-D Ising L = 64 periodic, 21 temperatures).

The metric's second half, tau_int-normalised effective samples/s (skip
with --no-ess): `ess_plain` for the plain sampler on the benchmark batch
itself (E and |M|, 400 samples of 5 sweeps), and `ess` on the 2-D skyrmion
Hamiltonian (L = 512 periodic, J = 1, Neel D = 0.3, h_z = 0.06,
T in {0.1, 0.2, 0.3}) for plain checkerboard sweeps and for sweeps
interleaved 4:1 with learned 16 x 16 block rounds -- vMF block proposals and
texture swaps (one network per temperature trained on 1e5 blocks of
plain-sampler configurations).

Prints one JSON line.
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

HBM_PEAK = 8.0e12


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Read + write bytes per second of device-to-device copies of a 1 GiB
    buffer (larger than the 256 MB Infinity Cache, so HBM-resident): the
    practical ceiling next to the 8 TB/s peak."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    secs = e0.elapsed_time(e1) / 1e3
    del a, b
    return 2.0 * nbytes * reps / secs


def cpu_baseline(budget_s=12.0):
    """The reference algorithm (oracle port) on this host: random-site
    single-site Metropolis over AoS double spins and vector<vector<int>>
    neighbour lists, mt19937, uniform-sphere trials (Old/lib/state.cpp:171-203)
    on the same 64^3 periodic lattice at T = 1.4, h = 0."""
    from oracle import oracle as orc
    orc.build()
    from tests.fixtures import box_locs
    L = 64
    d, J, D = orc.nn_bonds(3)
    f = orc.Field(box_locs(3, L), 3, L, 1, d, J, D)
    threads = max(1, min(16, os.cpu_count() or 1))
    probe, _ = f.bench_reference(False, 0.0, 1 / 1.4, 1, 1_000_000)
    attempts = int(max(1e6, probe * budget_s / 1.3))
    rate, secs = f.bench_reference(False, 0.0, 1 / 1.4, threads, attempts)
    return {"value": rate, "unit": "spin-updates/s", "cores": threads, "kind": "port",
            "sample": "%d threads x %d attempts (one chain per thread) on the 64^3 periodic Heisenberg "
                      "lattice, T=1.4, h=0, uniform-sphere trials; %.1f s" % (threads, attempts, secs)}


def plain_ess(lat, nsamples=400, samp_sweeps=5, extra_equil=1000):
    """tau_int-normalised effective samples/s of the plain sampler on the
    benchmark batch itself: after `extra_equil` more sweeps, `nsamples`
    samples of samp_sweeps sweeps + measurement (one replayed HIP graph);
    ESS/s = sum over chains of n / (2 tau_int) / wall time of the sampling."""
    from nanomagmc_amd.stats import ess
    lat.equil(extra_equil)
    lat.sync()
    lat.series_reset(nsamples, with_charge=False)
    t0 = time.perf_counter()
    lat.run_samples(nsamples, samp_sweeps)
    lat.sync()
    wall = time.perf_counter() - t0
    obs, _, _ = lat.get_series()
    out = {"samples": nsamples, "sweeps_per_sample": samp_sweeps, "chains": lat.ninst, "wall_s": wall}
    for key, series in (("E", obs[:, :, 10]), ("absM", np.linalg.norm(obs[:, :, 0:3], axis=-1))):
        n_eff, taus, closed = 0.0, [], 0
        for i in range(lat.ninst):
            ne, tau, ok = ess(series[i])
            n_eff += ne
            taus.append(tau)
            closed += int(ok)
        out[key] = {"ess_per_s": n_eff / wall, "tau_int_median": float(np.median(taus)),
                    "tau_int_max": float(np.max(taus)), "windows_closed": closed}
    return out


def ess_leg():
    """Effective samples/s of E and |Q| (Q_wrap, periodic plaquette sum):
    plain sweeps, sweeps + learned vMF block proposals, and sweeps + learned
    texture swaps; tau_int in samples of 4 sweeps each, Sokal window c = 6
    (`window_closed` false: the series is shorter than the window, tau is a
    lower bound and ESS an upper bound).  q_changes_per_ksweep counts the
    samples whose integer charge differs from the previous one: a frozen
    charge has a small tau_int without exploring anything."""
    from nanomagmc_amd.learn import compare_ess
    methods = ("vmf", "swap")
    r = compare_ess(L=512, temps=(0.1, 0.2, 0.3), h=0.06, D=0.3, replicas=2, anneal_sweeps=3000,
                    collect_rounds=100, collect_blocks=1000, train_steps=2000, batch=256, nsamples=2000,
                    methods=methods)
    rows = []
    for row in r["temps"]:
        def pick(m, k):
            st = row[m][k]
            return {"tau_int": st["tau_int"], "ess_per_s": st["ess_per_s"], "window_closed": st["window_closed"],
                    "mean": st["mean"], "err": st["err"]}
        rows.append({"T": row["T"], "train_loss": row["loss"], "block_acceptance": row["acceptance"],
                     **{m: {"E": pick(m, "E"), "absQ": pick(m, "absQ"),
                            "q_changes_per_ksweep": row[m]["q_changes_per_ksweep"]} for m in ("plain",) + methods},
                     "gain_absQ": row["gain_absQ"], "gain_E": row["gain_E"]})
    return {"metric": "tau_int-normalised effective samples/s", "unit": "samples/s",
            "config": "2-D skyrmion L=512 periodic, J=1, Neel D=0.3, h_z=0.06, %d replicas per method and T, "
                      "%d samples x 4 sweeps, learned: 1 round of 256 16x16 blocks per 4 sweeps "
                      "(swap: %d relaxation steps, core radius %.1f)"
                      % (r["replicas"], r["nsamples"], r["relax_iters"], r["core_radius"]),
            "train": {"blocks": r["train_blocks"], "steps": r["train_steps"], "batch": r["batch"],
                      "seconds": r["train_s"]},
            "s_per_sample": r["s_per_sample"], "temps": rows}


def learn_profile():
    """The workload of the learned-kernel profile: a 512 x 512 skyrmion
    lattice pair (J = 1, Neel D = 0.3, h = 0.06), 2e4 collected blocks, 50
    Adam steps at batch 256, then 20 rounds each of vMF block moves and
    texture swaps on both lattices.  Prints its wall time."""
    from nanomagmc_amd import Lattices
    from nanomagmc_amd.learn import Net, block_moves, block_swaps, neel_bonds
    t0 = time.perf_counter()
    lat = Lattices(2, False, True, neel_bonds(0.3), [{"n": (512, 512), "edge": 512}] * 2, use_j=True,
                   use_d=True, seed=3)
    lat.set_params([5.0, 5.0], 0.06)
    lat.randomize()
    lat.equil(200)
    net = Net(seed=1)
    for k in range(20):
        lat.equil(5)
        net.collect(lat, 0, 1000, seed=k)
    net.train(50, batch=256, lr=1e-3, seed=2)
    for k in range(20):
        block_moves(lat, net, [0, 1], k % 32, (3 * k) % 32)
        block_swaps(lat, net, [0, 1], k % 32, (5 * k) % 32)
    lat.sync()
    print(json.dumps({"learn_profile": "2 x 512^2 skyrmion lattices, 2e4 blocks, 50 Adam steps, 20 vMF + 20 swap "
                                       "rounds", "wall_s": time.perf_counter() - t0}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="heis3d", choices=["grains", "heis3d", "ising2d"])
    ap.add_argument("--trial", default=None, choices=["cone", "sphere"])
    ap.add_argument("--tune", action="store_true", help="autotune the sweep path before timing")
    ap.add_argument("--path", default=None, choices=["auto", "stream"], help="force a sweep path")
    ap.add_argument("--replicas", type=int, default=1,
                    help="heis3d: independent copies of every chain (working sets beyond the 256 MB cache)")
    ap.add_argument("--grain-sites", default=None, metavar="LO:HI",
                    help="grains: keep only grains with LO..HI present sites")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ess", action="store_true", help="skip the effective-samples leg")
    ap.add_argument("--learn-profile", action="store_true",
                    help="only run the learned-proposal workload profiled in profiles/learned_kernels_stats.md")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (reference for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.learn_profile:
        learn_profile()
        return
    if args.copy_calibration:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs independent chains per GPU and "
              "multi-GPU runs are not launched (use --gpus 1).", file=sys.stderr)
        sys.exit(2)
    import torch
    trial = args.trial or ("cone" if args.workload == "heis3d" else "sphere")
    t_setup = time.perf_counter()
    extra = {}
    if args.grain_sites:
        if args.workload != "grains":
            sys.exit("--grain-sites applies to --workload grains")
        lo, hi = (int(x) for x in args.grain_sites.split(":"))
        extra["site_range"] = (lo, hi)
    from nanomagmc_amd.workloads import WORKLOADS
    lat, sites, bpu, desc = WORKLOADS[args.workload](trial, replicas=args.replicas, **extra)
    setup_s = time.perf_counter() - t_setup
    total_sites = int(np.sum(sites))
    tuned = None
    if args.tune:
        from nanomagmc_amd.autotune import tune
        choice, timings = tune(lat, "%s/%s/%d" % (args.workload, trial, lat.ninst))
        tuned = {"choice": list(choice), "seconds_per_candidate": timings}
    if args.path:
        lat.set_path(args.path)
    # warm-up; cone half-angles adapt toward 50% acceptance every 10 sweeps
    # (on the device), then freeze
    if trial == "cone":
        lat.set_cone(np.full(lat.ninst, 1.0, dtype=np.float32))
        lat.equil_adapt(args.warmup, 10)
    else:
        lat.equil(args.warmup)
    lat.reset_accepts()
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(lat.stream())
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    lat.equil(args.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    acc = lat.accepts().astype(np.float64).sum() / (args.steps * total_sites)
    updates = args.steps * total_sites
    value = updates / wall
    # dominant kernel: one colour pass when streaming (2 launches per sweep)
    launches = 2 * args.steps
    launch_s = gpu_s / launches
    achieved = (total_sites / 2) * bpu / launch_s
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tpath) and args.replicas == 1 and not extra:   # measured for the default batch only
        tj = json.load(open(tpath))
        if tj.get("path", "auto") == (args.path or "auto"):   # and on the sweep path it was measured on
            traffic = tj.get("hbm_bytes_per_launch")
    out = {
        "metric": "Metropolis spin-updates/s and \u03c4_int-normalised effective samples/s, 1 GPU",
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if bpu == 36 else "int8",
        "data": "synthetic (spins from Philox, seeds 0x4E4D43xx)",
        "config": {"workload": desc, "instances": lat.ninst, "sites": total_sites, "acceptance": acc,
                   "setup_s": setup_s, "tuned": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": ("sweep colour pass (%d updates, %d B/update algorithmic)" % (total_sites // 2, bpu)
                                if args.workload != "grains" or args.path == "stream" else
                                "grain sweeps: compacted LDS-resident grains + streamed large grains; achieved = "
                                "%d B/update algorithmic x updates / time (half a sweep per unit)" % bpu),
                     "avg_launch_us": launch_s * 1e6,
                     "copy_GBps": copy_bandwidth() / 1e9},
    }
    if not args.no_ess:
        if args.workload == "heis3d" and args.replicas == 1:
            out["ess_plain"] = plain_ess(lat)
        out["ess"] = ess_leg()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
