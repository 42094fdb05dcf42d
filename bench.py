"""Throughput of the Metropolis hot path on one MI355X.

Default workload (the largest single-GPU configuration): 512 Weibull grains
(LATTSHAPE x, open boundaries, WEIBULLFACT 8) whose radius scale is
log-normal with mean 10 and sd 2.5 (truncated to [4, 20]), classical
Heisenberg spins with J = 1 and Bloch DMI D = 0.2, h_z in {0, 0.02, 0.05,
0.1} x 16 temperatures 2.0 -> 0.1 = 32768 lattice instances of uneven size
(179.2 M sites) resident in HBM, uniform-sphere trials (the reference's
move).  One step = one checkerboard sweep of every instance; the K timed
steps are one nmc_sweep call.

Other workloads (--workload): `heis3d` (3-D cubic Heisenberg L = 64
periodic, 32 T x 2 h = 64 instances, cone trials), `ising2d` (2-D Ising
L = 64 periodic, 21 temperatures).

The JSON line also carries:
* roofline: the kernel family that takes the most time (each family timed
  alone on its share of the batch with HIP events on the context's stream),
  bounded by VALU issue for the LDS-resident kernels and by HBM for the
  streaming ones, with its HBM traffic and VALU issue slots from
  rocprofv3 counter runs of the same workload (child processes started
  before this process touches the GPU, calibrated by a copy kernel of known
  size); every family under `parts`, the whole step under `step`;
* ess_plain: tau_int-normalised effective samples/s of E and |M| of the
  plain sampler on the batch itself; ess: the same for |Q| and E on the 2-D
  skyrmion lattice, plain against learned block proposals;
* cpu_baseline: the reference algorithm (the oracle's restatement, built
  -O3 -march=native on this host) on one grain of the batch, one core and
  all usable cores.
"""
import argparse
import json
import math
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
# wave64 VALU issue slots per second: 256 CUs x 4 SIMDs, each issuing one
# wave64 instruction per 4 cycles (16 lanes) at 2.4 GHz.  An fp32
# transcendental (v_sin, v_cos, v_sqrt, v_exp, ...) takes two slots: on this
# GPU a stream of them issues at half the rate of v_fma_f32 or v_mad_u64_u32
VALU_PEAK = 256 * 4 * 2.4e9 / 4
# algorithmic bytes per update: a spin is read as a target, read as a
# neighbour and written once per sweep (3 x 12 B fp32), plus one presence bit
# per colour pass for shaped lattices; Ising int8 3 B
BYTES = {"grains": 36.25, "heis3d": 36.0, "ising2d": 3.0}
# kernel families (nmc_sweep_kinds order): kernel-name prefix, bound
FAMILIES = [("k_sweep_heis<", "hbm"), ("k_sweep_heis_q<", "hbm"), ("k_sweep_span<", "hbm"), ("k_sweep_lds<", "valu"),
            ("k_sweep_lds_c<", "valu"), ("k_sweep_cstream<", "hbm")]
ISING_FAMILY = ("k_sweep_ising<", "hbm")
LDS_CHUNK = 256   # sweeps per LDS-resident launch


# ------------------------------------------------------------------ counters
def _run_pmc(counters, child, outdir, timeout=180):
    """One rocprofv3 counter pass over `bench.py <child>`; returns
    [(kernel name, {counter: value summed over dimensions})] per dispatch."""
    cmd = ["rocprofv3", "--pmc", *counters, "--output-format", "csv", "-d", outdir, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__), *child]
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                         start_new_session=True)
    try:
        rc = p.wait(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        raise RuntimeError("rocprofv3 pass %s timed out" % counters)
    path = None
    for dp, _, fs in os.walk(outdir):
        for f in fs:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(dp, f)
    if path is None:
        raise RuntimeError("rocprofv3 pass %s wrote no counters (rc %d)" % (counters, rc))
    import csv
    disp = {}
    for r in csv.DictReader(open(path)):
        d = disp.setdefault(r["Dispatch_Id"], (r["Kernel_Name"], {}))
        d[1][r["Counter_Name"]] = d[1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return [disp[k] for k in sorted(disp, key=int)]


def _per_kernel(dispatches, prefix, counter):
    return [c[counter] for name, c in dispatches if name.split("(")[0].split("nmc::")[-1].startswith(prefix)
            and counter in c]


def run_tuner(args):
    """The autotuner as a child process (before this one touches the GPU):
    returns (choice, seconds per candidate) or (None, error)."""
    cmd = [sys.executable, os.path.abspath(__file__), "--tune-only", *workload_args(args)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, start_new_session=True)
        d = json.loads(r.stdout.strip().splitlines()[-1])
        return tuple(d["choice"]), d["timings"]
    except (subprocess.TimeoutExpired, ValueError, IndexError, KeyError, OSError) as e:
        return None, "tuner failed: %s" % e


def workload_args(args):
    a = ["--workload", args.workload]
    if args.trial:
        a += ["--trial", args.trial]
    if args.replicas != 1:
        a += ["--replicas", str(args.replicas)]
    if args.grain_sites:
        a += ["--grain-sites", args.grain_sites]
    return a


def measure_counters(args, launch, steps=64, warmup=32):
    """HBM bytes (FETCH_SIZE + WRITE_SIZE, each scaled by what it reads for
    device copies of 256 MiB) and VALU issue slots (instructions, one more
    per fp32 transcendental) per dispatch of every
    sweep kernel family, from counter runs of this workload with `steps`
    timed sweeps after `warmup`.  For the LDS-resident kernels (one dispatch
    per size class and call) the per-sweep slope and the per-call intercept
    come from the warm-up and timed calls.  Both calls are long enough (32
    sweeps and more) to launch the size classes of the timed call; shorter
    calls merge the upper classes into one launch."""
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    # --families-apart: the counters are sampled per dispatch over the whole
    # device, so the kernel families (which share the GPU on two streams in
    # the timed run) are swept one after the other here
    child = workload_args(args) + ["--steps", str(steps), "--warmup", str(warmup), "--no-ess", "--no-cpu-baseline",
                                   "--no-counters", "--no-tune", "--no-parts", "--families-apart",
                                   "--launch-config", "%s:%d:%d" % tuple(launch)]
    tmp = tempfile.mkdtemp(prefix="nmc_pmc_")
    try:
        runs = {}
        for name, ctrs, ch in (("fetch", ["FETCH_SIZE"], child), ("write", ["WRITE_SIZE"], child),
                               ("valu", ["SQ_INSTS_VALU", "SQ_INSTS_VALU_TRANS_F32", "GRBM_GUI_ACTIVE"], child),
                               ("cfetch", ["FETCH_SIZE"], ["--copy-calibration"]),
                               ("cwrite", ["WRITE_SIZE"], ["--copy-calibration"])):
            runs[name] = _run_pmc(ctrs, ch, os.path.join(tmp, name))
    except (RuntimeError, OSError) as e:
        return None, str(e)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    copy_kib = 256 * 1024.0   # bytes copied per calibration dispatch / 1024
    cf = _per_kernel(runs["cfetch"], "__amd_rocclr_copyBuffer", "FETCH_SIZE")
    cw = _per_kernel(runs["cwrite"], "__amd_rocclr_copyBuffer", "WRITE_SIZE")
    if not cf or not cw:
        return None, "no copy kernel in the calibration runs"
    ffac, wfac = float(np.median(cf)) / copy_kib, float(np.median(cw)) / copy_kib
    fam = FAMILIES + [ISING_FAMILY]
    out = {"calibration": {"fetch_factor": ffac, "write_factor": wfac, "copy_bytes": int(copy_kib * 1024)},
           "steps": steps, "warmup": warmup, "kernels": {}}
    for prefix, _ in fam:
        f = _per_kernel(runs["fetch"], prefix, "FETCH_SIZE")
        w = _per_kernel(runs["write"], prefix, "WRITE_SIZE")
        v = _per_kernel(runs["valu"], prefix, "SQ_INSTS_VALU")
        tr = _per_kernel(runs["valu"], prefix, "SQ_INSTS_VALU_TRANS_F32")
        # issue slots: every instruction one, a transcendental one more
        if len(tr) == len(v):
            v = [a + b for a, b in zip(v, tr)]
        if not f:
            continue
        hbm = [1024.0 * (a / ffac + b / wfac) for a, b in zip(f, w)]
        rec = {"dispatches": len(f)}
        if prefix.startswith("k_sweep_lds"):
            # the warm-up call, then the timed call, each one dispatch per
            # size class of the family: per-call sums give the per-sweep
            # slope and the per-call intercept (loading and storing spins)
            half = len(hbm) // 2
            if half == 0 or len(hbm) != 2 * half:
                continue
            rec["dispatches_per_call"] = half
            ds = steps - warmup
            lo, hi = sum(hbm[:half]), sum(hbm[half:])
            # the per-sweep bytes are a small difference of two large
            # per-call figures (loading and storing the spins, and the first
            # call's cold tables): noise can make it negative
            rec["hbm_bytes_timed_call"] = hi
            rec["hbm_bytes_per_sweep"] = max(0.0, (hi - lo) / ds)
            if len(v) == len(f):
                vlo, vhi = sum(v[:half]), sum(v[half:])
                rec["valu_per_sweep"] = (vhi - vlo) / ds
                rec["valu_per_call"] = vlo - warmup * rec["valu_per_sweep"]
        else:
            rec["hbm_bytes_per_dispatch"] = float(np.mean(hbm))
            if v:
                rec["valu_per_dispatch"] = float(np.mean(v))
        out["kernels"][prefix.rstrip("<")] = rec
    return out, None


# ------------------------------------------------------------------ helpers
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


def timed_sweeps(lat, steps, inst=None):
    """Seconds of `steps` sweeps of the listed instances, from HIP events on
    the context's own stream (every sweep kernel is ordered on it)."""
    import torch
    stream = torch.cuda.ExternalStream(lat.stream())
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    lat.equil(0, inst=inst)      # builds the list's work tables before the first event
    lat.sync()
    ev0.record(stream)
    lat.equil(steps, inst=inst)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / 1e3


def usable_cores():
    """Host threads this job may use: the CPU affinity, capped by the cgroup
    quota and by OMP_NUM_THREADS (the job's CPU share where it is set)."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    use = info["affinity"]
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_quota_cpus"] = int(q) / int(per)
            use = min(use, max(1, int(int(q) // int(per))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        info["omp_num_threads"] = int(omp)
        use = min(use, int(omp))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["cpu_model"] = model
    return use, info


def cpu_baseline(workload, budget_s=8.0):
    """The reference algorithm on this host: random-site single-site
    Metropolis over AoS double spins and vector<vector<int>> neighbour lists,
    dE through nested std::function closures, mt19937 uniforms, uniform-sphere
    trials, log-form acceptance (Old/lib/state.cpp:171-203) -- the oracle's
    restatement, compiled -O3 -march=native here.  One independent chain per
    thread on one lattice of the workload: for the grain batch its median
    grain (J = 1, Bloch D = 0.2, T = 1, h = 0.05).  Runs on one core, then on
    every usable core."""
    from oracle import oracle as orc
    from tests.fixtures import box_locs
    orc.build()
    if workload == "grains":
        from nanomagmc_amd.workloads import grain_masks
        _, masks = grain_masks()
        nk = masks.sum(axis=1)
        g = int(np.argsort(nk)[len(nk) // 2])
        edge = 45
        locs = np.zeros((int(nk[g]), 3), dtype=np.int32)
        locs[:, :3] = np.argwhere(masks[g].reshape(edge, edge, edge))
        d, J, _ = orc.nn_bonds(3)
        f = orc.Field(locs, 3, edge, 0, d, J, 0.2 * d.astype(np.float64))
        ising, use_d, h, beta = False, True, 0.05, 1.0
        what = "grain %d of the batch (%d sites, open, J=1, Bloch D=0.2), T=1, h=0.05" % (g, nk[g])
    elif workload == "heis3d":
        d, J, D = orc.nn_bonds(3)
        f = orc.Field(box_locs(3, 64), 3, 64, 1, d, J, D)
        ising, use_d, h, beta = False, False, 0.0, 1 / 1.4
        what = "the 64^3 periodic Heisenberg lattice, T=1.4, h=0"
    else:
        d = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0]], dtype=np.int32)
        f = orc.Field(box_locs(2, 64), 2, 64, 1, d, np.ones(4), np.zeros((4, 3)))
        ising, use_d, h, beta = True, False, 0.0, 1 / 2.3
        what = "the 64^2 periodic Ising lattice, T=2.3"
    tmp = tempfile.mkdtemp(prefix="nmc_cpu_")
    try:
        fn = orc.bench_lib(tmp)
        probe, _ = orc.bench_reference(fn, f, ising, h, beta, 1, 500_000, use_d=use_d)
        n1 = int(max(1e6, probe * budget_s))
        r1, s1 = orc.bench_reference(fn, f, ising, h, beta, 1, n1, use_d=use_d)
        cores, info = usable_cores()
        rn, sn = orc.bench_reference(fn, f, ising, h, beta, cores, n1, use_d=use_d)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {"value": rn, "unit": "spin-updates/s", "cores": cores, "kind": "port",
            "single_core": r1, "host": info, "build": "g++ -O3 -march=native (oracle/Makefile benchref)",
            "sample": "%s, uniform-sphere trials: 1 thread x %d attempts (%.1f s), then %d threads x %d attempts "
                      "(one chain per thread, %.1f s)" % (what, n1, s1, cores, n1, sn)}


def plain_ess(lat, nsamples=400, samp_sweeps=5):
    """tau_int-normalised effective samples/s of E_s and |M| of the plain
    sampler on the batch itself: `nsamples` samples of samp_sweeps sweeps +
    measurement in one replayed HIP graph, ESS/s = sum over chains of
    n / (2 tau_int) / wall time of the sampling."""
    from nanomagmc_amd.stats import tau_int_batch
    lat.sync()
    lat.series_reset(nsamples, with_charge=False)
    t0 = time.perf_counter()
    lat.run_samples(nsamples, samp_sweeps)
    lat.sync()
    wall = time.perf_counter() - t0
    obs, _, _ = lat.get_series()
    out = {"samples": nsamples, "sweeps_per_sample": samp_sweeps, "chains": lat.ninst, "wall_s": wall}
    for key, series in (("E", obs[:, :, 10]), ("absM", np.linalg.norm(obs[:, :, 0:3], axis=-1))):
        tau, closed = tau_int_batch(series)
        out[key] = {"ess_per_s": float(np.sum(nsamples / (2.0 * tau)) / wall),
                    "tau_int_median": float(np.median(tau)), "tau_int_max": float(np.max(tau)),
                    "windows_closed": int(closed.sum())}
    return out


def ess_leg():
    """Effective samples/s of E and |Q| (Q_wrap, periodic plaquette sum):
    plain sweeps, sweeps + learned vMF block proposals, and sweeps + learned
    texture swaps, from starts every method shares (anneal, then swap-sampler
    preparation of every chain); tau_int in samples of 4 sweeps each, over
    each method's replicas together with the multi-chain autocorrelation
    (chains stuck at different values are charged), Sokal window c = 6
    (`window_closed` false: the series is shorter than the window, tau is a
    lower bound and ESS an upper bound).  q_changes_per_ksweep counts the
    samples whose integer charge differs from the previous one."""
    from nanomagmc_amd.learn import compare_ess
    methods = ("vmf", "swap")
    r = compare_ess(L=512, temps=(0.1, 0.2, 0.3), h=0.06, D=0.3, replicas=4, anneal_sweeps=3000,
                    collect_rounds=10000, collect_blocks=10, train_steps=2000, batch=256, eq_samples=10000,
                    nsamples=4000, methods=methods)
    rows = []
    for row in r["temps"]:
        def pick(m, k):
            st = row[m][k]
            return {"tau_int": st["tau_int"], "tau_int_per_chain": st["tau_int_per_chain"],
                    "ess_per_s": st["ess_per_s"], "window_closed": st["window_closed"], "mean": st["mean"],
                    "err": st["err"]}
        rows.append({"T": row["T"], "train_loss": row["loss"], "block_acceptance": row["acceptance"],
                     **{m: {"E": pick(m, "E"), "absQ": pick(m, "absQ"),
                            "q_changes_per_ksweep": row[m]["q_changes_per_ksweep"]} for m in ("plain",) + methods},
                     "gain_absQ": row["gain_absQ"], "gain_absQ_err": row["gain_absQ_err"],
                     "gain_E": row["gain_E"], "gain_E_err": row["gain_E_err"]})
    return {"metric": "tau_int-normalised effective samples/s", "unit": "samples/s",
            "config": "2-D skyrmion L=512 periodic, J=1, Neel D=0.3, h_z=0.06, %d replicas per method and T, "
                      "%d swap samples of preparation, then %d samples x 4 sweeps; learned: 1 round of 256 16x16 "
                      "blocks per lattice per 4 sweeps (swap: %d relaxation steps, core radius %.1f)"
                      % (r["replicas"], r["eq_samples"], r["nsamples"], r["relax_iters"], r["core_radius"]),
            "train": {"blocks": r["train_blocks"], "configurations": r["train_configurations"],
                      "steps": r["train_steps"], "batch": r["batch"], "seconds": r["train_s"],
                      "swap_loss": r["swap_loss"]},
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
        block_swaps(lat, net, [0, 1], k % 32, (5 * k) % 32, relax_iters=40)
    lat.sync()
    print(json.dumps({"learn_profile": "2 x 512^2 skyrmion lattices, 2e4 blocks, 50 Adam steps, 20 vMF + 20 swap "
                                       "rounds (40 relaxation steps)", "wall_s": time.perf_counter() - t0}), flush=True)


# ---------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="grains", choices=["grains", "heis3d", "ising2d"])
    ap.add_argument("--trial", default=None, choices=["cone", "sphere"])
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--tune-only", action="store_true", help="only run the autotuner and print its choice")
    ap.add_argument("--path", default=None, choices=["auto", "stream", "cstream"], help="force a sweep path")
    ap.add_argument("--launch-config", default=None, metavar="PATH:LDS:STREAM",
                    help="apply a launch shape (path, LDS workgroup size, streaming tile work-items)")
    ap.add_argument("--replicas", type=int, default=1,
                    help="heis3d: independent copies of every chain (working sets beyond the 256 MB cache)")
    ap.add_argument("--grain-sites", default=None, metavar="LO:HI",
                    help="grains: keep only grains with LO..HI present sites")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ess", action="store_true", help="skip the effective-samples legs")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 counter runs")
    ap.add_argument("--no-parts", action="store_true", help="skip timing each kernel family alone")
    ap.add_argument("--families-apart", action="store_true",
                    help="only sweep each kernel family alone (warm-up, then the steps), one family after the "
                         "other: the workload of the counter runs")
    ap.add_argument("--learn-profile", action="store_true",
                    help="only run the learned-proposal workload profiled in profiles/learned_kernels_stats.md")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (reference for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs independent chains per GPU and runs on more than "
              "one GPU are not launched (use --gpus 1).", file=sys.stderr)
        sys.exit(2)
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
    # the tuner and the counter runs first: child processes, started before
    # this one touches the GPU; all of them and this run use one launch shape
    launch, tuned = ("auto", 0, 0), None
    if args.launch_config:
        p, b, sb = args.launch_config.split(":")
        launch = (p, int(b), int(sb))
    elif args.path:
        launch = (args.path, 0, 0)
    elif not args.no_tune and not args.tune_only:
        choice, timings = run_tuner(args)
        if choice is not None:
            launch = choice
            tuned = {"choice": list(choice), "seconds_per_candidate": timings}
        else:
            tuned = {"error": timings}
    counters, counters_err = (None, "skipped (--no-counters)")
    if not args.no_counters and not args.tune_only:
        counters, counters_err = measure_counters(args, launch)
    import torch
    from nanomagmc_amd._lib import KIND_NAMES
    from nanomagmc_amd.workloads import WORKLOADS
    trial = args.trial or ("cone" if args.workload == "heis3d" else "sphere")
    t_setup = time.perf_counter()
    extra = {}
    if args.grain_sites:
        if args.workload != "grains":
            sys.exit("--grain-sites applies to --workload grains")
        lo, hi = (int(x) for x in args.grain_sites.split(":"))
        extra["site_range"] = (lo, hi)
    lat, sites, _, desc = WORKLOADS[args.workload](trial, replicas=args.replicas, **extra)
    bpu = BYTES[args.workload]
    sites = np.asarray(sites, dtype=np.int64)
    setup_s = time.perf_counter() - t_setup
    total_sites = int(sites.sum())
    from nanomagmc_amd import autotune
    if args.tune_only:
        choice, timings = autotune.tune(lat, "%s/%s/%d" % (args.workload, trial, lat.ninst), force=True)
        print(json.dumps({"choice": list(choice), "timings": timings}), flush=True)
        return
    autotune.apply(lat, launch)
    if args.families_apart:
        kinds = lat.sweep_kinds()
        for k in sorted(set(kinds.tolist())):
            ids = np.nonzero(kinds == k)[0]
            lat.equil(args.warmup, inst=ids)
            lat.equil(args.steps, inst=ids)
            lat.sync()
        return
    # warm-up; cone half-angles adapt toward 50% acceptance every 10 sweeps
    # (on the device), then freeze
    if trial == "cone":
        lat.set_cone(np.full(lat.ninst, 1.0, dtype=np.float32))
        lat.equil_adapt(args.warmup, 10)
    else:
        lat.equil(args.warmup)
    lat.reset_accepts()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu_s = timed_sweeps(lat, args.steps)
    wall = time.perf_counter() - t0
    acc = lat.accepts().astype(np.float64).sum() / (args.steps * total_sites)
    updates = args.steps * total_sites
    value = updates / wall
    step_alg = total_sites * bpu
    # every kernel family alone on its share of the batch, same sweep count
    kinds = lat.sweep_kinds()
    parts = []
    fams = FAMILIES if not lat.ising else [ISING_FAMILY] + FAMILIES[1:]
    for k in sorted(set(kinds.tolist())):
        ids = np.nonzero(kinds == k)[0]
        prefix, bound = fams[k]
        part = {"family": KIND_NAMES[k], "kernel": prefix.rstrip("<"), "bound": bound, "instances": int(ids.size),
                "sites": int(sites[ids].sum())}
        if not args.no_parts and len(set(kinds.tolist())) > 1:
            part["seconds"] = timed_sweeps(lat, args.steps, inst=ids)
        elif not args.no_parts:
            part["seconds"] = gpu_s
        parts.append(part)
    kc = (counters or {}).get("kernels", {})
    kc_steps = (counters or {}).get("steps")
    for part in parts:
        if "seconds" not in part:
            continue
        per_sweep = part["seconds"] / args.steps
        alg = part["sites"] * bpu                          # algorithmic bytes per sweep of this share
        part["ms_per_sweep"] = per_sweep * 1e3
        part["alg_GBps"] = alg / per_sweep / 1e9
        c = kc.get(part["kernel"], {})
        if part["family"] in ("lds", "lds_compact"):
            # a launch here is one chunk of up to LDS_CHUNK sweeps: one
            # dispatch per size class, back to back on one stream
            launches = math.ceil(args.steps / LDS_CHUNK)
            part["launches"] = launches
            if "dispatches_per_call" in c:
                part["dispatches_per_launch"] = c["dispatches_per_call"]
            part["avg_launch_us"] = part["seconds"] / launches * 1e6
            sweeps_per_launch = args.steps / launches
            if "hbm_bytes_per_sweep" in c:
                steps_c = kc_steps if kc_steps else sweeps_per_launch
                part["traffic"] = c["hbm_bytes_timed_call"] + c["hbm_bytes_per_sweep"] * (sweeps_per_launch - steps_c)
            if "valu_per_sweep" in c:
                part["valu_per_launch"] = c["valu_per_call"] + c["valu_per_sweep"] * sweeps_per_launch
                part["valu_Gslots_per_s"] = part["valu_per_launch"] / (part["avg_launch_us"] * 1e-6) / 1e9
                part["valu_frac"] = part["valu_Gslots_per_s"] * 1e9 / VALU_PEAK
        else:
            # one colour pass per launch; the time includes, spread over the
            # sweeps of the call, any per-call copies of the family
            part["launches"] = 2 * args.steps
            part["avg_launch_us"] = part["seconds"] / (2 * args.steps) * 1e6
            part["alg_bytes_per_launch"] = alg / 2
            part["hbm_frac"] = part["alg_bytes_per_launch"] / (part["avg_launch_us"] * 1e-6) / HBM_PEAK
            if "hbm_bytes_per_dispatch" in c:
                part["traffic"] = c["hbm_bytes_per_dispatch"]
            if "valu_per_dispatch" in c:
                part["valu_per_launch"] = c["valu_per_dispatch"]
                part["valu_Gslots_per_s"] = part["valu_per_launch"] / (part["avg_launch_us"] * 1e-6) / 1e9
                part["valu_frac"] = part["valu_Gslots_per_s"] * 1e9 / VALU_PEAK
            # a streaming pass is bound by whichever of HBM and VALU issue it
            # uses the larger share of
            if part.get("valu_frac", 0.0) > part["hbm_frac"]:
                part["bound"] = "valu"
    timed = [p for p in parts if "seconds" in p]
    dom = max(timed, key=lambda p: p["seconds"]) if timed else None
    roof = {"step": {"alg_bytes": step_alg, "achieved": step_alg / (gpu_s / args.steps) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": step_alg / (gpu_s / args.steps) / HBM_PEAK,
                     "note": "algorithmic bytes of a whole sweep of the batch / sweep time (all families)"},
            "copy_GBps": copy_bandwidth() / 1e9, "parts": parts,
            "counters": counters if counters else {"error": counters_err}}
    if dom is not None:
        if dom["bound"] == "valu":
            ach = dom.get("valu_Gslots_per_s")
            if dom["family"] in ("lds", "lds_compact"):
                what = ("%s: %d instances, %d sites, %d sweeps per launch (%s dispatches, one per size class, back "
                        "to back)" % (dom["kernel"], dom["instances"], dom["sites"],
                                      round(args.steps / dom["launches"]), dom.get("dispatches_per_launch", "?")))
            else:
                what = ("%s: one colour pass of %d instances (%d updates)"
                        % (dom["kernel"], dom["instances"], dom["sites"] // 2))
            roof.update({"bound": "valu", "achieved": ach, "peak": VALU_PEAK / 1e9, "unit": "G issue slots/s",
                         "frac": None if ach is None else ach * 1e9 / VALU_PEAK,
                         "traffic": dom.get("traffic"),
                         "kernel": what + "; achieved = wave64 VALU issue slots per launch (instructions, one more "
                                          "per fp32 transcendental; counters, the family swept alone) / launch "
                                          "time (events, the family alone)",
                         "avg_launch_us": dom["avg_launch_us"]})
        else:
            ach = dom["alg_bytes_per_launch"] / (dom["avg_launch_us"] * 1e-6)
            roof.update({"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": ach / HBM_PEAK, "traffic": dom.get("traffic"),
                         "kernel": "%s: one colour pass of %d instances (%d updates, %.2f B/update algorithmic)"
                                   % (dom["kernel"], dom["instances"], dom["sites"] // 2, bpu),
                         "avg_launch_us": dom["avg_launch_us"]})
    out = {
        "metric": "Metropolis spin-updates/s and τ_int-normalised effective samples/s, 1 GPU",
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8" if lat.ising else "f32",
        "data": "synthetic (spins from Philox, seeds 0x4E4D43xx)",
        "config": {"workload": desc, "instances": lat.ninst, "sites": total_sites, "acceptance": acc,
                   "setup_s": setup_s, "tuned": tuned, "launch": list(launch),
                   "gpu_ms_per_step": gpu_s / args.steps * 1e3},
        "roofline": roof,
    }
    if not args.no_ess:
        out["ess_plain"] = plain_ess(lat)
        out["ess"] = ess_leg()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
