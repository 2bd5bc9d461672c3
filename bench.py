"""Throughput of the fused RHS + SSP-RK3 step on one MI355X.

Metric (value): cell updates per second of the batched FP32 form -- forced
HIT, 64 instances, 16 each at 128^2, 256^2, 384^2 and 512^2, seeds 0-63,
Mach (0.1, 0.25, 0.4)[seed % 3], production settings otherwise (viscosity
7.5e-4, OU forcing in 3-5 with tau 0.5, Mach control, drag 0.10 for k <= 2,
cooling 5.0, hyperviscosity every 5 steps), FP32 state with FP64 reductions,
autotuned launch shapes, one batch step (every instance one full SSP-RK3
step with dt, Mach control, forcing noise and hyperviscosity) per HIP graph
replay, exactly --steps replays timed.  The same line carries the FP64 side
of the comparison (production HIT 512^2 FP64, one instance on device
control: the plain FP64 form), Riemann configuration 3 at 512^2 FP64 (the
single-plan driver), the FP32-vs-FP64 error after 1000 steps, a serving
stream with p50/p99 time to solution, the roofline of the dominant kernel
and the CPU baseline.

Roofline: the stage update (b_update, the largest kernel of the batched
step): algorithmic bytes per launch (stage 0 reads the stage state, dF/dx,
dG/dy and the forcing and writes the new state, 18 W per node; stages 1-2
also read q0, 22 W) over the average launch time from HIP events on the
batch's stream around the three stage updates launched back to back five
times after the timed replays;
traffic from rocprofv3 FETCH_SIZE and WRITE_SIZE counter runs of the same
launch shapes (child processes), scaled by what the counters read for a
512 MiB device copy.

The parent process never touches the GPU: the measurement runs in a child
process (so the counter runs can be started afterwards), the CPU baseline
(the NumPy oracle on the host) in the parent.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import csv
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mcell-updates/s per RK3 step, FP64 vs FP32 batched; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-update of one full RK3 step (dependency-minimal count)
HIT_STEP_BYTES = {"fp64": 635.0, "fp32": 320.0}
RIEMANN_STEP_BYTES = {"fp64": 614.0, "fp32": 310.0}
COPY_MB = 512
PROBE_STEPS = 5


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--phase", choices=("main", "gpu", "probe"), default="main", help=argparse.SUPPRESS)
    ap.add_argument("--tiles", type=int, nargs=4, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--kernel-steps", type=int, default=6, help="eager steps with HIP events around b_update")
    ap.add_argument("--no-extras", action="store_true", help="only the batched FP32 measurement")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=12)
    ap.add_argument("--serving-rate", type=float, default=2.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--serving-steps", type=int, nargs=2, default=(200, 5000), metavar=("MIN", "MAX"))
    ap.add_argument("--no-fp32-error", action="store_true")
    ap.add_argument("--fp32-steps", type=int, default=1000)
    ap.add_argument("--save-states", default=None, help="directory for the final FP32/FP64 HIT 512^2 states")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=16, help="process-pool size of the CPU baseline (at most)")
    return ap.parse_args()


# ================================================================ GPU child
def _events_time(fn, k):
    import torch

    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(k):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return max(time.perf_counter() - t0, e0.elapsed_time(e1) / 1e3)


def batched_fp32(args):
    """The metric: the 64-instance FP32 batch, autotuned, K timed graph replays."""
    import torch

    from hit_solver_amd.batch import BatchRunner, hit_batched_configs

    cs = hit_batched_configs("fp32")
    cells = sum(c.nx * c.ny for c in cs)
    b = BatchRunner(cs)
    for _ in range(2):
        b.step()
    best, _times = b.autotune()
    progress(f"autotuned launch shapes {best}")
    for _ in range(max(0, args.warmup - args.kernel_steps)):
        b.step()
    # b_update time from HIP events on the batch's stream (eager steps)
    b.timing(True)
    for _ in range(args.kernel_steps):
        b.step()
    upd_ms, upd_n = b.timing_read()
    b.timing(False)
    b.capture(1)
    b.replay(1)
    torch.cuda.synchronize()
    T = _events_time(lambda: b.replay(1), args.steps)
    value = cells * args.steps / T / 1e6
    W = 4
    N = cells
    # per-stage algorithmic bytes of the stage update: stage 0 (q0 is the stage state) 18 W, stages 1-2 22 W
    upd_alg = (18 * W + 22 * W + 22 * W) / 3.0 * N
    # the kernel's own duration: after the timed replays, the three stage
    # updates launched back to back between two HIP events on the batch's
    # stream (the event pairs around each update of the eager steps above
    # also hold the wait for the stage's sweeps on the other stream)
    upd_ms_one, upd_bb = b.time_updates(reps=5)
    upd_avg_s = upd_ms_one / 1e3
    res = {"value": value, "T": T, "cells": cells, "instances": len(cs), "tiles": list(best),
           "upd_avg_us": upd_avg_s * 1e6, "upd_launches": upd_bb, "upd_alg_bytes": upd_alg,
           "upd_achieved_GBs": upd_alg / upd_avg_s / 1e9,
           "upd_eager_event_us": upd_ms / max(upd_n, 1) * 1e3, "upd_eager_launches": upd_n}
    del b
    torch.cuda.empty_cache()
    return res


def hit_fp64_plain(steps):
    """Production HIT 512^2 FP64, one instance on device control (the plain
    FP64 GPU form), and through the Python operator API."""
    import torch

    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver

    cfg = HIT2DConfig.production_512(dtype="fp64", tfinal=1e9)
    b = BatchRunner([cfg])
    for _ in range(5):
        b.step()
    b.capture(10)
    b.replay(1)
    reps = max(2, steps // 10)
    T = _events_time(lambda: b.replay(1), reps)
    plain = 512 * 512 * reps * 10 / T / 1e6
    del b
    s = HITSolver(cfg)
    for _ in range(3):
        s.step()
    k = max(10, steps // 4)
    Ta = _events_time(s.step, k)
    del s
    torch.cuda.empty_cache()
    return {"hit512_fp64_plain": {"value": round(plain, 3), "unit": "Mcell-updates/s",
                                  "ms_per_step": round(T / (reps * 10) * 1e3, 4),
                                  "form": "one instance, device-side control, 10 steps per HIP graph"},
            "hit512_fp64_api": {"value": round(512 * 512 * k / Ta / 1e6, 3), "unit": "Mcell-updates/s",
                                "ms_per_step": round(Ta / k * 1e3, 4),
                                "form": "HITSolver.step: Python operator API, host-side dt/Mach/forcing control"}}


def riemann_c3(steps):
    """Riemann configuration 3, 512^2 FP64, single plan, graph replay."""
    from hit_solver_amd.riemann import RiemannConfig3, RiemannSolver

    cfg = RiemannConfig3(nx=512, ny=512, cfl=0.4, tfinal=0.3, sensor_width=4, jump_threshold=0.025,
                         shear_threshold=None, mn=0.001, hyperviscosity_interval=1, progress_every=10**9)
    s = RiemannSolver(cfg)
    for _ in range(10):
        s.step(validate=False)
    s.capture()
    T = _events_time(lambda: s.replay(1), steps)
    v = 512 * 512 * steps / T / 1e6
    return {"value": round(v, 3), "unit": "Mcell-updates/s", "ms_per_step": round(T / steps * 1e3, 4),
            "step_roofline_frac": round(v * 1e6 * RIEMANN_STEP_BYTES["fp64"] / 1e9 / HBM_PEAK_GBS, 4),
            "form": "single plan, one step (3 fused stages + hyperviscosity) per HIP graph",
            "workload": "Lax-Liu configuration 3, 512x512 FP64 hybrid, CFL 0.4, mn 0.001 every step, guard 2"}


def fp32_error(steps, save_dir=None):
    """Production HIT 512^2 (seed 1234): FP32 batched form vs FP64, per field
    relative L2 at checkpoints."""
    import dataclasses

    import numpy as np
    import torch

    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig

    c64 = HIT2DConfig.production_512(dtype="fp64", tfinal=1e9, max_steps=steps)
    c32 = dataclasses.replace(c64, dtype="fp32")
    a, b = BatchRunner([c64]), BatchRunner([c32])
    b.tune(32, 64, 2, 3)
    a.capture(10)
    b.capture(10)
    names = ("rho", "rho_u", "rho_v", "E")
    out = {"workload": f"forced HIT 512x512 production settings, seed 1234, {steps} steps; FP32 batched form vs "
                       "FP64 form on the device", "steps": steps, "checkpoints": {}}
    done = 0
    for ck in sorted({min(steps, x) for x in (10, 100, 500, steps)}):
        reps = (ck - done) // 10
        a.replay(reps)
        b.replay(reps)
        done += reps * 10
        torch.cuda.synchronize()
        q64, q32 = a.q[0].double(), b.q[0].double()
        out["checkpoints"][str(done)] = {nm: float(f"{float(torch.linalg.vector_norm(q32[k] - q64[k]) / torch.linalg.vector_norm(q64[k])):.3e}")
                                         for k, nm in enumerate(names)}
    out["rel_l2_final"] = out["checkpoints"][str(done)]
    out["max_rel_l2"] = max(out["rel_l2_final"].values())
    if save_dir:
        os.makedirs(save_dir, exist_ok=True)
        np.save(os.path.join(save_dir, f"hit512_fp32_step{done}.npy"), b.q[0].cpu().numpy())
        np.save(os.path.join(save_dir, f"hit512_fp64_step{done}.npy"), a.q[0].cpu().numpy())
    del a, b
    torch.cuda.empty_cache()
    return out


def serving(count, rate, steps):
    from hit_solver_amd.serving import ServingEngine, poisson_stream

    reqs = poisson_stream(count, rate, seed=2024, steps=tuple(steps))
    eng = ServingEngine(round_steps=10)
    _res, summary = eng.run(reqs)
    cells = sum(r.n * r.n * r.steps for r in reqs)
    summary.update({"unit_latency": "s", "arrival_rate_per_s": rate, "seed": 2024,
                    "stream": f"{count} requests, kinds uniform over c3/c3_08/c6/hit, n uniform in 128..1024 step "
                              f"64, fp32/fp64 uniform, {steps[0]}-{steps[1]} steps each, {cells / 1e9:.2f} G "
                              "cell-updates in total",
                    "form": "continuous batching: one batch per dtype (HIT and Riemann solves mixed) on its own "
                            "stream, rounds of 10 steps as HIP graphs, HBM-budget admission (80% of free device "
                            "memory); phases = host wall seconds by phase"})
    for k in ("p50_s", "p99_s", "mcell_updates_per_s"):
        if summary.get(k) is not None:
            summary[k] = round(summary[k], 4)
    return summary


def gpu_phase(args):
    import torch

    torch.cuda.set_device(0)
    out = {"batched": batched_fp32(args)}
    progress(f"batched FP32 {out['batched']['value']:.1f} Mcell-updates/s")
    if not args.no_extras:
        out["hit"] = hit_fp64_plain(max(60, args.steps))
        progress("plain FP64 HIT done")
        out["riemann_c3_512_fp64"] = riemann_c3(max(50, args.steps))
        progress("Riemann C3 done")
        if not args.no_fp32_error:
            out["fp32_error"] = fp32_error(args.fp32_steps, args.save_states)
            progress("FP32 error run done")
        if not args.no_serving:
            out["serving"] = serving(args.serving_requests, args.serving_rate, args.serving_steps)
            progress("serving stream done")
    print(json.dumps(out), flush=True)


# ================================================================ counter probe
def probe_phase(args):
    """A few eager steps of the batch with fixed launch shapes (run under
    rocprofv3 --pmc by the parent), after a 512 MiB device copy for the
    counter calibration."""
    import torch

    from hit_solver_amd.batch import BatchRunner, hit_batched_configs

    torch.cuda.set_device(0)
    src = torch.ones(COPY_MB << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    b = BatchRunner(hit_batched_configs("fp32"))
    b.tune(*args.tiles)
    for _ in range(PROBE_STEPS):  # hyperviscosity every 5 steps: one pass in the window
        b.step()
    torch.cuda.synchronize()
    print("probe done", flush=True)


def counter_run(counter, tiles, workdir):
    """Per-dispatch counter values of the probe (kernel name -> list)."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", str(workdir), "-o", "run", "--",
           sys.executable, str(ROOT / "bench.py"), "--phase", "probe", "--tiles", *map(str, tiles)]
    env = dict(os.environ, TMPDIR="/tmp", PYTHONPATH=str(ROOT))
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=240)
    files = list(Path(workdir).rglob("*counter_collection.csv"))
    if not files:
        raise RuntimeError(f"no counter file (rc {r.returncode}): {r.stderr[-400:]}")
    vals: dict = {}
    for row in csv.DictReader(open(files[0])):
        if row.get("Counter_Name") != counter:
            continue
        key = (row["Dispatch_Id"], row["Kernel_Name"])
        vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
    by_kernel: dict = {}
    for (_d, name), v in vals.items():
        by_kernel.setdefault(name, []).append(v)
    return by_kernel


def traffic(tiles):
    """HBM bytes per b_update launch (FETCH_SIZE and WRITE_SIZE runs, KiB
    scaled by the 512 MiB copy's readings) and per batch step."""
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            progress(f"rocprofv3 --pmc {counter}")
            bk = counter_run(counter, tiles, Path(td) / counter)
            # the three 512 MiB copies are the largest copy dispatches (the set-up makes small ones)
            copy = sorted(v for n, vs in bk.items() if "copyBuffer" in n for v in vs)[-3:]
            scale = (COPY_MB * 1024.0) / (sum(copy) / len(copy))  # KiB the copy must move / KiB read
            upd = [v for n, vs in bk.items() if "b_update" in n for v in vs]
            # per step: every batched kernel of the probe's steps (not the copies or the set-up)
            step_kib = sum(v for n, vs in bk.items() if "hitx::b_" in n for v in vs) / PROBE_STEPS
            out[counter] = {"scale": scale, "upd_bytes": sum(upd) / len(upd) * 1024.0 * scale,
                            "step_bytes": step_kib * 1024.0 * scale}
    return {"upd_bytes": out["FETCH_SIZE"]["upd_bytes"] + out["WRITE_SIZE"]["upd_bytes"],
            "step_bytes": out["FETCH_SIZE"]["step_bytes"] + out["WRITE_SIZE"]["step_bytes"],
            "scale_fetch": round(out["FETCH_SIZE"]["scale"], 4), "scale_write": round(out["WRITE_SIZE"]["scale"], 4)}


# ================================================================ CPU baseline
def _pool_worker(job):
    """Oracle steps of a few batch instances on one core: per instance one
    warm-up step, then `steps` timed steps; returns (cell-updates, seconds)."""
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import hit as OH

    cells, secs = 0, 0.0
    for kw, steps in job:
        run = OH.HITRun(OH.HITParams(**kw))
        run.step()
        t0 = time.perf_counter()
        for _ in range(steps):
            run.step()
        secs += time.perf_counter() - t0
        cells += kw["nx"] * kw["ny"] * steps
    return cells, secs


def _instance_params(c):
    import dataclasses

    from oracle import hit as OH

    names = {f.name for f in dataclasses.fields(OH.HITParams)}
    return {k: v for k, v in dataclasses.asdict(c).items() if k in names}


def cpu_baseline(workers):
    """The NumPy oracle on the host: the 64-instance workload in a process
    pool (aggregate cell-updates/s), and one 512^2 FP64 production instance
    on one core (median step over 20 steps after 2 warm-ups)."""
    import multiprocessing as mp
    import statistics

    from hit_solver_amd.batch import hit_batched_configs
    from oracle import hit as OH

    nproc = len(os.sched_getaffinity(0))
    P = max(1, min(workers, nproc))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cfgs = hit_batched_configs("fp64")
    jobs = [[] for _ in range(P)]
    order = sorted(range(len(cfgs)), key=lambda i: -cfgs[i].nx)  # largest first, round robin
    for j, i in enumerate(order):
        jobs[j % P].append((_instance_params(cfgs[i]), 2))
    progress(f"CPU baseline: process pool of {P} on {nproc} cores")
    with mp.get_context("spawn").Pool(P) as pool:
        res = pool.map(_pool_worker, jobs)
    cells = sum(r[0] for r in res)
    busiest = max(r[1] for r in res)
    pool_v = cells / busiest / 1e6
    progress("CPU baseline: single core 512^2")
    from hit_solver_amd.hit2d import HIT2DConfig

    os.environ["OMP_NUM_THREADS"] = "1"
    run = OH.HITRun(OH.HITParams(**_instance_params(HIT2DConfig.production_512(dtype="fp64", tfinal=1e9))))
    for _ in range(2):
        run.step()
    ts = []
    for _ in range(20):
        t0 = time.perf_counter()
        run.step()
        ts.append(time.perf_counter() - t0)
    med = statistics.median(ts)
    return {"value": round(pool_v, 4), "unit": "Mcell-updates/s", "cores": P, "kind": "port",
            "sample": f"NumPy oracle (the reference's algorithm, FP64) on the 64-instance workload: a pool of {P} "
                      f"single-threaded processes, each instance 1 warm-up + 2 timed steps, {cells / 1e6:.1f} M "
                      f"cell-updates over the busiest worker's {busiest:.1f} s",
            "host": {"os_cpu_count": os.cpu_count(), "nproc": nproc, "model": model},
            "single_core_hit512_fp64": {"value": round(512 * 512 / med / 1e6, 4), "unit": "Mcell-updates/s",
                                        "median_step_s": round(med, 4),
                                        "sample": "production HIT 512^2, median of 20 steps after 2 warm-ups, 1 core"}}


# ================================================================ main
def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "hit_solver_amd runs on exactly one GPU; instances are independent, so more "
                                   "GPUs would each take a share of the instance table (not built here)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.phase == "gpu":
        return gpu_phase(args)
    if args.phase == "probe":
        return probe_phase(args)
    child = [sys.executable, str(ROOT / "bench.py"), "--phase", "gpu", "--steps", str(args.steps), "--warmup",
             str(args.warmup), "--kernel-steps", str(args.kernel_steps), "--fp32-steps", str(args.fp32_steps),
             "--serving-requests", str(args.serving_requests), "--serving-rate", str(args.serving_rate),
             "--serving-steps", *map(str, args.serving_steps)]
    for flag in ("no_extras", "no_serving", "no_fp32_error"):
        if getattr(args, flag):
            child.append("--" + flag.replace("_", "-"))
    if args.save_states:
        child += ["--save-states", args.save_states]
    r = subprocess.run(child, stdout=subprocess.PIPE, text=True)
    if r.returncode != 0:
        sys.exit(r.returncode)
    g = json.loads(r.stdout.strip().splitlines()[-1])
    bt = g["batched"]
    value = bt["value"]
    roof = {"bound": "hbm", "kernel": "hitx::b_update<float> (stage update of the batched step)",
            "achieved": round(bt["upd_achieved_GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bt["upd_achieved_GBs"] / HBM_PEAK_GBS, 4), "traffic": None,
            "alg_bytes_per_launch": round(bt["upd_alg_bytes"]), "avg_launch_us": round(bt["upd_avg_us"], 2),
            "launches_timed": bt["upd_launches"],
            "timing": "HIP events on the batch's stream around the three stage updates launched back to back 5 "
                      "times after the timed replays (launch-to-launch gaps included)",
            "eager_event_us": round(bt["upd_eager_event_us"], 2),
            "eager_event_note": f"event pairs around each update of {args.kernel_steps} eager steps, which also "
                                "hold the wait for the stage's sweeps on the other stream"}
    if not args.no_traffic:
        try:
            tr = traffic(bt["tiles"])
            roof["traffic"] = round(tr["upd_bytes"])
            roof["traffic_step_bytes"] = round(tr["step_bytes"])
            roof["traffic_alg_step_bytes"] = round(bt["cells"] * HIT_STEP_BYTES["fp32"])
            roof["traffic_scale"] = {"FETCH_SIZE": tr["scale_fetch"], "WRITE_SIZE": tr["scale_write"]}
            roof["traffic_source"] = (f"rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (child runs, {PROBE_STEPS} eager "
                                      "steps of the same launch shapes), scaled by the counters' readings for a "
                                      f"{COPY_MB} MiB device copy")
        except Exception as e:  # counters unavailable: traffic stays null
            roof["traffic_error"] = str(e)[:300]
    step_bytes = HIT_STEP_BYTES["fp32"]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bt["T"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (random-phase streamfunction initial states, seeds 0-63; device forcing noise)",
        "config": {"workload": "batched forced HIT: 64 instances, 16 each at 128^2, 256^2, 384^2, 512^2, FP32 state "
                               "with FP64 reductions, production settings, autotuned launch shapes",
                   "instances": bt["instances"], "cells": bt["cells"], "tiles": bt["tiles"],
                   "graph": "one batch step per HIP graph replay"},
        "roofline": roof,
        "step_roofline": {"alg_bytes_per_cell_update": step_bytes,
                          "achieved_GBs": round(value * 1e6 * step_bytes / 1e9, 1),
                          "frac": round(value * 1e6 * step_bytes / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if "hit" in g:
        out["fp64"] = g["hit"]
        out["fp32_batched_vs_fp64_plain"] = round(value / g["hit"]["hit512_fp64_plain"]["value"], 3)
    for k in ("riemann_c3_512_fp64", "fp32_error", "serving"):
        if k in g:
            out[k] = g[k]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_workers)
        progress("CPU baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
