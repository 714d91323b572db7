#!/usr/bin/env python3
"""Throughput of the pseudospectral step on one MI355X.

Workload (the headline "grid-pt*steps/s at 2048^2 FP32"): one instance on a
2048x2048 grid, FP32 transforms and updates with FP64 shift/time bookkeeping,
Lx = Ly = 2*pi, sigma = 1, nu = kappa = 1e-4, g = rho0 = rho0' = 1,
delt = 1e-4, cfl number 0.35; omega = 0 and a seeded random-phase rho on
2 <= |k| <= 32 with amplitude 1e-4 |k|^-1.5.  Synthetic data; the state lives
in HBM before the timed region starts.

A step is one time_advance: three kernels (or, fused, P2 and the kernel that
runs P3 of a step with P1 of the next), replayed from a captured HIP graph.  The timed region is `--steps` steps between HIP events on the
library's stream, after `--warmup` untimed steps.  Per-kernel durations come
from a second, kernel-by-kernel pass in which every dispatch records HIP
events at its own start and end (with the P3+P1 fusion, which the library
uses for captured batches at 2048^2 FP32, a third pass times the fused kernel
the same way); the roofline object is for the kernel of the timed steps with
the largest average duration (P1, P2, P3; fused: P2 and P3+P1).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0
METRIC = "grid-pt*steps/s at 2048^2 FP32 + HBM roofline %; time-to-error learned vs plain"


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(s, c):
    """Algorithmic HBM bytes per launch of the three step kernels.

    P1 reads omega^n, rho^n (2M) and writes six y-transformed fields (6I);
    P2 reads those (6I) and writes two bracket spectra (2I); P3 reads them (2I)
    plus omega^n, rho^n, omega^{n-1}, rho^{n-1} and four tendency histories (8M)
    and writes omega^{n+1}, rho^{n+1} and two tendencies (4M).  c = bytes per
    complex element.
    """
    M, I = s["M"], s["I"]
    return {"p1": c * (2 * M + 6 * I), "p2": c * 8 * I, "p3": c * (2 * I + 12 * M)}


def fused_bytes(s, c):
    """Algorithmic HBM bytes per launch of the fused kernel (P3 of step n with
    P1 of step n+1): P3's c(2I + 12M) plus P1's six G fields (6I); P1's read of
    the new state is not made, the rows stay on chip."""
    return c * (8 * s["I"] + 12 * s["M"])


def step_bytes(s, c, fused, G):
    """Algorithmic HBM bytes per step of the kernels a captured batch of G
    steps launches: P1, P2, P3 per step, or fused one P1, G P2, G-1 fused
    P3+P1 and one P3."""
    kb = kernel_bytes(s, c)
    if not fused or G < 2:
        return sum(kb.values())
    return kb["p2"] + ((G - 1) * fused_bytes(s, c) + kb["p1"] + kb["p3"]) / G


def read_traffic(name):
    """HBM bytes per launch of a step kernel as two rocprofv3 --pmc runs of this
    bench measured them (FETCH_SIZE, WRITE_SIZE, calibrated by a copy of known
    size; profiles/traffic.json.txt, written by --pmc-summary), with the
    profiled kernel's name: (bytes, source) or (None, None) if not recorded."""
    path = os.path.join(REPO, "profiles", "traffic.json.txt")
    if not os.path.exists(path):
        return None, None
    try:
        d = json.load(open(path)).get(name)
        if d is None:
            return None, None
        return d["bytes"], f"profiles/traffic.json.txt: rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE of {d.get('kernel', name)}"
    except Exception:
        return None, None


def pmc_summary(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, cal_bytes):
    """HBM bytes per dispatch of each step kernel from rocprofv3 --pmc runs.

    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what they read
    for the plain copy kernel k_copy16 moving `cal_bytes` (read once, written
    once).  Returns {"p1_<n>_<fp32|fp64>": bytes, ..., "calibration": {...}}.
    """
    import collections
    import csv
    import re

    def per_kernel(path):
        acc = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}

    def copy_kib(d):
        v = [x for k, x in d.items() if "k_copy16" in k]
        if not v:
            raise RuntimeError("calibration run holds no k_copy16 dispatch")
        return v[0]

    f_cal = cal_bytes / (copy_kib(per_kernel(cal_fetch_csv)) * 1024.0)
    w_cal = cal_bytes / (copy_kib(per_kernel(cal_write_csv)) * 1024.0)
    fetch, write = per_kernel(fetch_csv), per_kernel(write_csv)
    out = {"calibration": {"copy_bytes": cal_bytes, "fetch_factor": f_cal, "write_factor": w_cal}}
    pat = re.compile(r"k_(p[123])(?:_row|_pair)?<(float|double), (\d+)[,>]")
    fused = re.compile(r"k_p3<(?:float|double), \d+, (?:true|false), \d+, true")   # P3 with the next P1
    for k, fv in fetch.items():
        m = pat.search(k)
        if not m or k not in write:
            continue
        name = "p31" if fused.search(k) else m.group(1)
        key = f"{name}_{m.group(3)}_{'fp32' if m.group(2) == 'float' else 'fp64'}"
        out[key] = {"bytes": f_cal * fv * 1024.0 + w_cal * write[k] * 1024.0,
                    "read_bytes": f_cal * fv * 1024.0, "write_bytes": w_cal * write[k] * 1024.0, "kernel": k}
    return out


def cpu_baseline(orc, n, params, ic, budget_s):
    """Time the FP64 oracle (the CPU restatement) on a bounded sample."""
    threads = max(1, min(16, os.cpu_count() or 1))
    o = orc.Oracle(threads=threads, nx=n, ny=n, **params)
    o.set_spectral("rho", orc.seeded_spectrum(n, n, params["Lx"], params["Ly"], **ic))
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    steps = int(max(2, min(1000, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": n * n * steps / dt, "unit": "grid-pt*steps/s", "cores": threads, "kind": "port",
            "sample": f"{n}^2 FP64 oracle, seeded IC, {steps} steps after 1 warm step, "
                      f"{threads} OpenMP threads, {dt:.2f} s"}


# Physics and initial condition of the timed workload.
PARAMS = dict(Lx=2 * np.pi, Ly=2 * np.pi, delt=1e-4, cfl_num=0.35, nu=1e-4, kappa=1e-4,
              sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0, rho_eps1=1e-2, noise=0, linear=0)
IC = dict(seed=20250101, kmin=2.0, kmax=32.0, amp=1e-4, expo=-1.5)


def cpu_scan(orc, params, ic, budget_s):
    """The FP64 oracle on the host cores at one thread and at all of them (at
    most 16): 128^2 with the small-grid physics for 2000 steps in full, 1024^2
    and 2048^2 with the bench physics for as many steps as fit in `budget_s`
    seconds.  One JSON line per (size, threads)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), model)
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    top = max(1, min(16, avail))
    small = dict(params, delt=1e-3, nu=1e-3, kappa=1e-3, sigma=0.5)
    small_ic = dict(seed=ic["seed"], kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    for n, full, p, q in ((128, 2000, small, small_ic), (1024, 0, params, ic), (2048, 0, params, ic)):
        for threads in sorted({1, top}):
            o = orc.Oracle(threads=threads, nx=n, ny=n, **p)
            o.set_spectral("rho", orc.seeded_spectrum(n, n, p["Lx"], p["Ly"], **q))
            t0 = time.perf_counter()
            o.step(1)
            one = time.perf_counter() - t0
            steps = full or int(max(2, min(1000, budget_s / max(one, 1e-6))))
            t0 = time.perf_counter()
            o.step(steps)
            dt = time.perf_counter() - t0
            o.close()
            print(json.dumps({"metric": "FP64 oracle throughput", "n": n, "threads": threads, "steps": steps,
                              "seconds": dt, "ms_per_step": dt * 1e3 / steps,
                              "grid_pt_steps_per_s": n * n * steps / dt, "cpu": model,
                              "cpus_available": avail}), flush=True)


def ensemble_members(spec, seed=7):
    """Parameters of the ensemble members: SPEC is NxCOUNT,... or mixed:COUNT
    (sizes drawn uniformly from 128, 256, 512).  Per instance: shear rate
    U(0, 2), g U(0.5, 2), nu = kappa log-uniform in [1e-4, 1e-3], initial delt
    2e-4 under CFL number 0.3 (check_cfl halves it as needed), rho seeded on
    2 <= |k| <= n/16 with its own seed."""
    rng = np.random.default_rng(seed)
    if spec.startswith("mixed:"):
        sizes = [int(x) for x in rng.choice([128, 256, 512], int(spec.split(":")[1]))]
    else:
        sizes = []
        for part in spec.split(","):
            n, cnt = (int(x) for x in part.split("x"))
            sizes += [n] * cnt
    out = []
    for n in sizes:
        nu = float(10 ** rng.uniform(-4, -3))
        prm = dict(nx=n, ny=n, Lx=2 * np.pi, Ly=2 * np.pi, delt=2e-4, cfl_num=0.3, nu=nu, kappa=nu,
                   sigma=float(rng.uniform(0.0, 2.0)), g=float(rng.uniform(0.5, 2)), rho0=1.0, rho0_prime=1.0,
                   rho_eps1=1e-2, noise=0, linear=0)
        ic = dict(seed=int(rng.integers(1 << 62)), kmin=2.0, kmax=max(2.0, n / 16), amp=1e-4, expo=-1.5)
        out.append((prm, ic))
    return out


# Variant choices (P1, P3, fusion) tried per grid size on an ensemble's batched
# group: P1 0 row, 1 four-row tile, 2 pair; P3 0 pair, 1 tile, 2 split.
ENSEMBLE_CANDIDATES = [(0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 1, 0), (1, 1, 0), (2, 1, 0), (0, 0, 1), (0, 1, 1)]


def tune_ensemble_group(shrt, group, graph_steps, steps=100):
    """The variant choice of one grid size for a batched ensemble: each
    candidate of ENSEMBLE_CANDIDATES times `steps` ensemble steps of scratch
    instances built from the group's (params, initial condition) list; returns
    (p1, p3, fuse) of the fastest and the times.  Single-instance tuning
    (shrt_autotune) optimises one instance's launch latency; a batched group of
    small grids is bound by its workgroups' lane occupancy instead (the tile
    variants give whole wavefronts below 512^2)."""
    times = {}
    for p1, p3, fuse in ENSEMBLE_CANDIDATES:
        gs = []
        for prm, ic in group:
            g = shrt.Shrt(precision=shrt.FP32, graph_steps=0, **prm)
            g.initialize_seeded(**ic)
            g.set_variants(p1, 0, p3)
            g.set_fusion(fuse)
            gs.append(g)
        e = shrt.Ensemble(gs, steps_per_graph=graph_steps)
        e.step(graph_steps)
        times[(p1, p3, fuse)] = e.time_steps(steps) / steps
        e.close()
        for g in gs:
            g.close()
    return min(times, key=times.get), times


def run_ensemble(shrt, args):
    """Aggregate throughput of independent instances stepped by one captured
    graph per batch (shrt_ensemble_*), members from ensemble_members.  The
    variants of each grid size are tuned on that size's batched group
    (tune_ensemble_group) and set on all its members.  The roofline fraction
    sums each member's algorithmic bytes per step over the ensemble's time per
    step."""
    members, points, sizes, nbytes, nbytes_c14 = [], 0, [], 0, 0
    specs = ensemble_members(args.ensemble)
    for prm, ic in specs:
        g = shrt.Shrt(precision=shrt.FP32, graph_steps=0, **prm)
        g.initialize_seeded(**ic)
        members.append(g)
        n = prm["nx"]
        sizes.append(n)
        points += n * n
    tuned = {}
    for n in sorted(set(sizes)):
        (p1, p3, fuse), times = tune_ensemble_group(shrt, [sp for sp, m in zip(specs, sizes) if m == n],
                                                    args.graph_steps)
        tuned[n] = dict(variants=[p1, 0, p3], fuse=fuse,
                        ms_per_step={f"{a},{b},{c}": round(t, 5) for (a, b, c), t in times.items()})
        for g, m in zip(members, sizes):
            if m == n:
                g.set_variants(p1, 0, p3)
                g.set_fusion(fuse)
    for n in sizes:
        s = shrt.sizes(n, n)
        nbytes += step_bytes(s, 8, tuned[n]["fuse"], args.graph_steps)
        nbytes_c14 += 8 * (14 * s["M"] + 10 * s["I"])
    log(f"ensemble: {len(members)} instances, per-size choices {tuned}")
    ens = shrt.Ensemble(members, steps_per_graph=args.graph_steps)
    ens.step(args.graph_steps)   # warm-up and capture
    steps = max(args.graph_steps, args.steps // args.graph_steps * args.graph_steps)
    ms = ens.time_steps(steps)
    out = {"metric": "ensemble grid-pt*steps/s", "value": points * steps / (ms * 1e-3),
           "unit": "grid-pt*steps/s", "instances": len(members), "spec": args.ensemble, "steps": steps,
           "ms_per_step": ms / steps, "dtype": "fp32", "graph_steps": args.graph_steps,
           "sizes": {str(n): sizes.count(n) for n in sorted(set(sizes))},
           "roofline": {"bound": "hbm", "bytes_per_step": nbytes, "unit": "GB/s",
                        "achieved": nbytes / (ms / steps * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                        "frac": nbytes / (ms / steps * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "bytes_per_step_c14M10I": nbytes_c14,
                        "frac_c14M10I": nbytes_c14 / (ms / steps * 1e-3) / 1e9 / PEAK_HBM_GBS},
           "tuned": {str(n): v for n, v in tuned.items()}}
    ens.close()
    for g in members:
        g.close()
    print(json.dumps(out), flush=True)


CLOSURE_DEFAULT = dict(n_fine=256, n_coarse=64, n_mid=128, delt=1e-3, interval=20, spinup=1000, snaps=40,
                       nu=1.25e-4, rho0_prime=-1.0, amp=0.5, kmin=1.0, kmax_ic=16.0, kmax_net=16, width=16, layers=4)

# The learned-closure configuration at full size: 2048^2 fine runs (shear U(0.5, 1.5),
# g = 1, nu = kappa = 1e-4, delt 1.25e-4 so that max|vx + sigma*y| stays under the
# CFL number) filtered to 256^2, 64 trajectories x 4000 steps (1000 of spin-up, then
# a snapshot every 20 steps), a 4-layer width-32 network on 32 modes per axis, 16
# held-out seeds; plain 256^2, 512^2, 1024^2 and 2048^2 runs give the
# error-versus-GPU-time curve the learned 256^2 run is placed on.  The correction
# strength (0 = plain) and band are chosen by 4000-step rollouts on the 4
# validation trajectories before the held-out seeds are run.
CLOSURE_CONFIG5 = dict(n_fine=2048, n_coarse=256, n_mid=512, delt=1.25e-4, interval=20, spinup=1000, snaps=150,
                       nu=1e-4, rho0_prime=-1.0, amp=2.0, kmin=1.0, kmax_ic=32.0, kmax_net=32, width=32, layers=4,
                       cfl=0.5, trajectories=64, val_trajectories=4, test_seeds=list(range(900, 916)),
                       eval_steps=4000, eval_grids=[512, 1024, 2048], host_path=False, iters=3000, batch=16,
                       kcut=32, guard=0.1, calib_alphas=[0.125, 0.25, 0.5, 1.0], calib_kcuts=[8, 16, 32])


def run_closure(shrt, args):
    """Learned closure versus plain solvers: data from fine runs, device
    training with early stopping, then held-out seeds measured against the
    filtered fine run (enstrophy spectrum and rho relative L2), with the GPU
    time of every run and the plain method's time at the learned run's error."""
    cl = _load("closure_pipe", os.path.join(REPO, "shearing-rt_amd", "closure.py"))
    spec = json.loads(args.closure)
    kw = dict(CLOSURE_CONFIG5 if spec.pop("config5", False) else CLOSURE_DEFAULT)
    kw.update(spec)
    ntrain, nval = kw.pop("trajectories", 64), kw.pop("val_trajectories", 4)
    iters, lr, batch = kw.pop("iters", 2000), kw.pop("lr", 5e-4), kw.pop("batch", 16)
    tests, evsteps = kw.pop("test_seeds", [900, 901, 902, 903]), kw.pop("eval_steps", 1000)
    grids, host_path = kw.pop("eval_grids", None), kw.pop("host_path", True)
    alphas, kcuts = kw.pop("calib_alphas", None), kw.pop("calib_kcuts", None)
    cfg = cl.Config(**kw)
    p = cl.Pipeline(shrt, cfg)
    t0 = time.perf_counter()
    x, y = p.generate(range(100, 100 + ntrain), log=log)
    val = p.generate(range(800, 800 + nval))
    t_data = time.perf_counter() - t0
    log(f"closure: {len(x)} training samples ({p.dropped} non-finite dropped, fine-run CFL halvings "
        f"{p.max_halvings}) in {t_data:.1f} s; input std "
        f"{np.round(x[:, :3].std(axis=(0, 2)), 4)}, target std {np.round(y.std(axis=(0, 2)), 4)}")
    t0 = time.perf_counter()
    losses = p.fit(x, y, iters=iters, batch=batch, lr=lr, val=val)
    t_train = time.perf_counter() - t0
    del x, y
    log(f"closure: trained {iters} iterations in {t_train:.1f} s, best validation R2 {p.val_best:.3f}")
    calib = None
    if alphas:
        # correction strength and band from rollouts on the validation seeds (alpha 0: plain)
        t0 = time.perf_counter()
        calib = p.calibrate(range(800, 800 + nval), evsteps, alphas=alphas, kcuts=kcuts, log=log)
        log(f"closure: calibrated alpha {cfg.alpha}, kcut {cfg.kcut} in {time.perf_counter() - t0:.1f} s")
    res = p.evaluate(tests, evsteps, grids=grids, host_path=host_path, log=log)
    runs = list(res[0]["wall_s"].keys())
    mean = lambda run, key: float(np.mean([r[run][key] for r in res]))
    wall = lambda run: float(np.mean([r["wall_s"][run] for r in res]))
    out = {"metric": "learned closure: enstrophy-spectrum / rho relative L2 error vs filtered fine run",
           "config": dict(kw, trajectories=ntrain, iters=iters, eval_steps=evsteps, test_seeds=tests,
                          eval_grids=grids),
           "samples": int(p.nsamples), "data_s": t_data, "train_s": t_train, "train_loss_last": float(losses[-1]),
           "val_r2_best": p.val_best, "val_curve": p.val_curve,
           "calibration": calib, "alpha": cfg.alpha, "kcut": cfg.kcut,
           "error": {run: {"enstrophy_spectrum": mean(run, "enstrophy_spectrum"), "rho": mean(run, "rho")}
                     for run in runs},
           "wall_s_per_run": {run: wall(run) for run in runs},
           "skipped_corrections": int(sum(r.get("skipped_device", 0) for r in res)), "per_seed": res}
    if host_path:
        out["device_vs_host_omega_rel_l2"] = float(max(r["device_vs_host"] for r in res))
    # time to error: the plain method's GPU time at the learned run's error, per measure
    plain_runs = [r for r in runs if r == "plain" or r.startswith("plain_") or r == "mid"]
    tte = {}
    for key in ("enstrophy_spectrum", "rho"):
        pts = [(wall(r), mean(r, key)) for r in plain_runs]
        t_eq, how = cl.time_at_error(pts, mean("closure_device", key))
        tte[key] = {"learned_s": wall("closure_device"), "learned_error": mean("closure_device", key),
                    "plain_s_at_equal_error": t_eq, "bracket": how,
                    "ratio_plain_over_learned": t_eq / wall("closure_device"),
                    "plain_points": [{"run": r, "s": wall(r), "error": mean(r, key)} for r in plain_runs]}
    out["time_to_error"] = tte
    p.net.close()
    print(json.dumps(out), flush=True)


def run_vendor_fft(shrt, args):
    """The reference's per-step transforms (six 2-D C2R, two R2C) by the vendor
    FFT library (shearing-rt_amd/vendor_fft_bench), beside this library's whole
    step at the same size and precision."""
    import subprocess
    exe = os.path.join(REPO, "shearing-rt_amd", "vendor_fft_bench")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build it with `make -C shearing-rt_amd`")
    n = args.n
    p = subprocess.run([exe, str(n), str(n), "100", args.precision], capture_output=True, text=True, timeout=300)
    if p.returncode != 0:
        raise RuntimeError(p.stderr)
    vendor = json.loads(p.stdout.strip().splitlines()[-1])
    prec = shrt.FP32 if args.precision == "fp32" else shrt.FP64
    g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=args.graph_steps, Lx=2 * np.pi, Ly=2 * np.pi, delt=1e-4,
                  cfl_num=0.35, nu=1e-4, kappa=1e-4, sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0, rho_eps1=1e-2,
                  noise=0, linear=0)
    g.initialize_seeded(seed=20250101, kmin=2.0, kmax=32.0, amp=1e-4, expo=-1.5)
    g.time_advance(args.graph_steps * 5)
    steps = max(args.graph_steps, 1000 // args.graph_steps * args.graph_steps)
    ms = g.time_steps(steps) / steps
    g.close()
    print(json.dumps({"metric": "vendor FFT comparison", "n": n, "precision": args.precision, "vendor": vendor,
                      "shrt_ms_per_step": ms,
                      "vendor_step_ffts_over_shrt_step": vendor["ms_reference_step_ffts"] / ms}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="steps per captured graph (100: 79.6 us/step at 2048^2 FP32 against 80.1 with 20)")
    ap.add_argument("--kernel-steps", type=int, default=50)
    ap.add_argument("--variants", default="", metavar="P1,P2,P3",
                    help="kernel variants of the three passes (shrt_set_variants); default: the library's")
    ap.add_argument("--fusion", type=int, choices=[0, 1], default=None,
                    help="captured batches fuse P3 of step n with P1 of step n+1 (1) or not (0); "
                         "default: the library's choice for the size and precision")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES (the HBM-counter reference) and exit")
    ap.add_argument("--pmc-summary", nargs=5, metavar=("FETCH_CSV", "WRITE_CSV", "CAL_FETCH_CSV",
                                                       "CAL_WRITE_CSV", "CAL_BYTES"),
                    help="only turn rocprofv3 FETCH_SIZE/WRITE_SIZE counter files into HBM bytes per "
                         "dispatch (JSON on stdout) and exit")
    ap.add_argument("--autotune-report", default="", metavar="SIZES",
                    help="only run shrt_autotune for each square size (e.g. 256,512,1024,2048,4096) and "
                         "print the per-variant kernel times as JSON lines, then exit")
    ap.add_argument("--vendor-fft", action="store_true",
                    help="only time the reference's per-step FFTs with the vendor library (comparison) and exit")
    ap.add_argument("--closure", default="", metavar="JSON",
                    help="only run the learned-closure experiment (JSON overrides of its configuration, e.g. '{}') "
                         "and exit")
    ap.add_argument("--cpu-scan", action="store_true",
                    help="only time the FP64 oracle on the host cores (1 and all threads; 128^2 x 2000 steps, "
                         "1024^2 and 2048^2 on --cpu-budget seconds each) and exit")
    ap.add_argument("--ensemble", default="", metavar="SPEC",
                    help="only time an ensemble, e.g. 128x24,256x24,512x16 (grid x count; FP32, mixed "
                         "shear rates and seeds, one captured graph per --graph-steps steps) and exit")
    args = ap.parse_args()

    if args.pmc_summary:
        f, w, cf, cw, nb = args.pmc_summary
        print(json.dumps(pmc_summary(f, w, cf, cw, int(nb)), indent=1), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} (WORLD_SIZE {world}) is not supported; "
              "the single-instance step is not sharded across GPUs", file=sys.stderr)
        sys.exit(2)

    orc = _load("shrt_oracle_py", os.path.join(REPO, "oracle", "oracle.py"))
    if args.cpu_scan:
        cpu_scan(orc, PARAMS, IC, args.cpu_budget)
        return
    shrt = _load("shrt_binding", os.path.join(REPO, "shearing-rt_amd", "shrt.py"))
    if shrt.lib().shrt_device_count() < 1:
        raise RuntimeError("no HIP device visible")

    def copy_gbs(nbytes, reps, stream=False):
        ms = ctypes.c_double()
        fn = shrt.lib().shrt_stream_copy_bench if stream else shrt.lib().shrt_copy_bench
        st = fn(nbytes, reps, ctypes.byref(ms))
        if st != 0:
            raise RuntimeError(shrt.lib().shrt_last_error().decode())
        return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9

    if args.copy_calibrate:
        ms, gbs = copy_gbs(args.copy_calibrate, 5)
        print(json.dumps({"copy_bytes": args.copy_calibrate, "ms": ms, "GB/s": gbs}), flush=True)
        return

    if args.ensemble:
        run_ensemble(shrt, args)
        return
    if args.closure:
        run_closure(shrt, args)
        return
    if args.vendor_fft:
        run_vendor_fft(shrt, args)
        return
    if args.autotune_report:
        prec = shrt.FP32 if args.precision == "fp32" else shrt.FP64
        for n in (int(x) for x in args.autotune_report.split(",")):
            g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=0, Lx=2 * np.pi, Ly=2 * np.pi, delt=1e-4,
                          cfl_num=0.35, nu=1e-4, kappa=1e-4, sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0,
                          rho_eps1=1e-2, noise=0, linear=0)
            t = g.autotune(reps=20)
            names = ["p1_row", "p1_tile", "p1_pair", "p2_three", "p2_four", "p3_pair", "p3_tile", "p3_split"]
            print(json.dumps({"n": n, "precision": args.precision, "choice": [t["p1"], t["p2"], t["p3"]],
                              "us_per_launch": {k: round(v * 1e3, 2) for k, v in zip(names, t["ms"])},
                              "fuse": t["fuse"], "us_per_step": {"unfused": round(t["step_ms"][0] * 1e3, 2),
                                                                 "fused": round(t["step_ms"][1] * 1e3, 2)}}),
                  flush=True)
            g.close()
        return

    n = args.n
    prec = shrt.FP32 if args.precision == "fp32" else shrt.FP64
    params = dict(PARAMS)
    ic = dict(IC)

    log(f"bench: {n}^2 {args.precision}, graph_steps={args.graph_steps}, warmup={args.warmup}, steps={args.steps}")
    g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=args.graph_steps, **params)
    if args.fusion is not None:
        g.set_fusion(args.fusion)
    if args.variants:
        g.set_variants(*[int(v) for v in args.variants.split(",")])
    used = g.variants()
    g.initialize_seeded(**ic)
    if args.warmup > 0:
        g.time_advance(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ms = g.time_steps(args.steps)
    wall = time.perf_counter() - t0
    st = g.state()
    log(f"bench: timed {args.steps} steps: {ms:.3f} ms (events), {wall * 1e3:.3f} ms (wall); state {st}")
    kt = g.kernel_times(args.kernel_steps)
    log(f"bench: kernel ms per launch p1={kt[0]:.4f} p2={kt[1]:.4f} p3={kt[2]:.4f}")
    kt31 = g.fused_kernel_time(args.kernel_steps) if used["fuse"] else None
    if kt31 is not None:
        log(f"bench: fused P3+P1 kernel ms per launch {kt31:.4f}")
    g.close()
    copy_ms, copy_bw = copy_gbs(1 << 30, 10)
    scopy_ms, scopy_bw = copy_gbs(1 << 30, 10, stream=True)
    log(f"bench: plain copy 1 GiB: {copy_ms:.3f} ms, {copy_bw:.1f} GB/s; streaming (nontemporal) copy: "
        f"{scopy_ms:.3f} ms, {scopy_bw:.1f} GB/s (read + write)")

    c = 8 if prec == shrt.FP32 else 16
    s = shrt.sizes(n, n)
    kb = kernel_bytes(s, c)
    names = ["p1", "p2", "p3"]
    kms = {k: float(v) for k, v in zip(names, kt)}
    # the kernels of the timed steps: P1, P2, P3, or with fusion P2 and the fused
    # P3+P1 (P1 and P3 then run once per captured batch)
    if kt31 is not None:
        kms["p31"] = float(kt31)
        kb_all = dict(kb, p31=fused_bytes(s, c))
        timed = ["p2", "p31"]
    else:
        kb_all = kb
        timed = names
    dname = max(timed, key=lambda k: kms[k])
    achieved = kb_all[dname] / (kms[dname] * 1e-3) / 1e9
    traffic, traffic_src = read_traffic(f"{dname}_{n}_{args.precision}")
    ms_step = ms / args.steps
    nbytes = step_bytes(s, c, kt31 is not None, min(args.graph_steps, args.steps))
    baseline_formula = c * (14 * s["M"] + 10 * s["I"])

    out = {
        "metric": METRIC,
        "value": n * n * args.steps / (ms * 1e-3),
        "unit": "grid-pt*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic: omega=0, seeded random-phase rho on 2<=|k|<=32, amplitude 1e-4|k|^-1.5",
        "config": {"workload": f"pseudospectral step, {n}^2 {args.precision.upper()} single instance "
                               "(sigma=1, nu=kappa=1e-4, g=1, delt=1e-4)",
                   "nx": n, "ny": n, "graph_steps": args.graph_steps, "fusion": used["fuse"],
                   "variants": f"{used['p1']},{used['p2']},{used['p3']}" + ("" if args.variants else " (default)")},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "bytes_per_launch": kb_all[dname], "ms_per_launch": kms[dname]},
        "kernels_ms": kms,
        "stream_copy_GBs": copy_bw,
        "stream_copy_nt_GBs": scopy_bw,
        "step_roofline": {"bytes_per_step": nbytes,
                          "frac": nbytes / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS,
                          "frac_of_stream_copy_nt": nbytes / (ms_step * 1e-3) / 1e9 / scopy_bw,
                          "bytes_per_step_c14M10I": baseline_formula,
                          "frac_c14M10I": baseline_formula / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "final_state": {"istep": st["istep"], "time": st["time"], "delt": st["delt"],
                        "remaps": st["remap_count"], "halvings": st["halvings"]},
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (FP64 oracle) ...")
        out["cpu_baseline"] = cpu_baseline(orc, n, params, ic, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
