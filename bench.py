"""Throughput of the fused RHS + SSP-RK3 step on one MI355X.

Headline (value): the single-instance FP64 Riemann workload below.  The same
JSON line also carries the forced-HIT measurements ("hit"): production HIT
512x512 FP64 as one instance (through the Python operator API with its host
control, and on device-side control replayed as a HIP graph), and the
batched FP32 form: 64 instances, 16 each at 128^2, 256^2, 384^2 and 512^2,
seeds 0-63, Mach (0.1, 0.25, 0.4)[seed % 3], production settings otherwise,
autotuned column-walk tiles, steps replayed as a HIP graph.

Default workload: Lax-Liu Riemann configuration 3, 512x512, FP64, hybrid
compact/WENO7 scheme, CFL 0.4, T 0.3 (fixed dt, 1343 steps for the full
solve), sensor width 4, jump threshold 0.025, shear term off, compact-node
hyperviscosity mn 0.001 every step, outflow guard 2 cells.  One "step" is a
full RK3 step (3 fused stages: reductions, sensor, x/y sweeps with the compact
solves, stage update and clean) plus the post-step hyperviscosity pass with
a fresh sensor mask.  The state lives in HBM before timing starts.

Prints one JSON line:  value = cell updates per second / 1e6 (whole job),
roofline of the dominant kernel (the line sweep: algorithmic bytes per
launch / average launch time from HIP events on its stream), and the
NumPy oracle timed on a bounded sample on the host as cpu_baseline.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell-update of one full RK3 step, Riemann path with
# hyperviscosity every step (dependency-minimal count: per stage one x-pass
# reading q and writing dF/dx, one y-pass reading q, dF/dx, q0 and writing q,
# the uint8 mask written and read once; plus the hyperviscosity pass).
STEP_BYTES = {"fp64": 614.0, "fp32": 310.0}
# forced HIT (forcing, drag, cooling, viscous terms; hyperviscosity every 5 steps)
HIT_STEP_BYTES = {"fp64": 635.0, "fp32": 320.0}
# HBM bytes per line-sweep launch (both sweeps of a stage) measured with rocprofv3 FETCH_SIZE and WRITE_SIZE
# (scaled by a 512 MiB copy kernel's readings; profiles/k_line_riemann_c3_512_fp64.md), for the workloads that
# were profiled
LINE_TRAFFIC = {(512, "fp64"): 34_454_851}


def progress(msg):
    """One line per finished phase on stderr (the result line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--kernel-steps", type=int, default=10, help="eager steps timed per sweep launch")
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--dtype", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the host oracle sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hit", action="store_true", help="skip the forced-HIT measurements")
    ap.add_argument("--hit-steps", type=int, default=60)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream measurement")
    ap.add_argument("--no-fp32-error", action="store_true", help="skip the FP32-vs-FP64 1000-step error")
    ap.add_argument("--fp32-steps", type=int, default=1000)
    ap.add_argument("--save-states", default=None, help="directory for the final FP32/FP64 HIT 512^2 states")
    ap.add_argument("--serving-requests", type=int, default=24)
    ap.add_argument("--serving-rate", type=float, default=4.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--serving-steps", type=int, nargs=2, default=(200, 1000), metavar=("MIN", "MAX"))
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run 3 plain device copies of this many MiB first (scale for rocprofv3 FETCH/WRITE_SIZE)")
    return ap.parse_args()


def cpu_baseline(nx, seconds):
    """The NumPy oracle (the reference's algorithm, one core) on a bounded
    sample of the same workload: Riemann config 3 at nx^2, steps from t=0."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import riemann as OR

    run = OR.RiemannRun(OR.config3(nx=nx, ny=nx, shear_threshold=None, progress_every=10**9))
    run.step()  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        run.step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": round(nx * nx * n / el / 1e6, 4), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"NumPy oracle, Riemann C3 {nx}x{nx} FP64, {n} RK3 steps (+HV) after 1 warm-up step, "
                      f"{el:.1f} s on 1 host core"}


def _timed(fn, k):
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


def hit_measurements(steps):
    """Forced HIT: plain single-instance FP64 (API path and device-control
    graph) and the batched FP32 form."""
    import torch

    from hit_solver_amd.batch import BatchRunner, hit_batched_configs
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver

    res = {}
    cfg = HIT2DConfig.production_512(dtype="fp64", tfinal=1e9)
    s = HITSolver(cfg)
    for _ in range(5):
        s.step()
    k = max(10, steps // 3)
    T = _timed(s.step, k)
    res["hit512_fp64_api"] = {"value": round(512 * 512 * k / T / 1e6, 3), "unit": "Mcell-updates/s",
                              "ms_per_step": round(T / k * 1e3, 4),
                              "form": "HITSolver.step: Python operator API, host-side dt/Mach/forcing control"}
    del s
    b = BatchRunner([cfg])
    for _ in range(5):
        b.step()
    b.capture(10)
    b.replay(1)
    T = _timed(lambda: b.replay(1), max(1, steps // 10))
    n = max(1, steps // 10) * 10
    plain = 512 * 512 * n / T / 1e6
    res["hit512_fp64_plain"] = {"value": round(plain, 3), "unit": "Mcell-updates/s", "ms_per_step": round(T / n * 1e3, 4),
                                "form": "one instance, device-side control, 10 steps per HIP graph"}
    del b
    torch.cuda.empty_cache()
    cs = hit_batched_configs("fp32")
    cells = sum(c.nx * c.ny for c in cs)
    b = BatchRunner(cs)
    for _ in range(2):
        b.step()
    best, _times = b.autotune()
    b.capture(5)
    b.replay(1)
    reps = max(1, steps // 5)
    T = _timed(lambda: b.replay(1), reps)
    n = reps * 5
    v = cells * n / T / 1e6
    ach = v * 1e6 * HIT_STEP_BYTES["fp32"] / 1e9
    res["hit_batched_fp32"] = {
        "value": round(v, 3), "unit": "Mcell-updates/s", "ms_per_step": round(T / n * 1e3, 4),
        "instances": len(cs), "cells": cells, "tiles": {"rb": best[0], "rbp": best[1], "line_form": ("block", "wave")[best[2] - 1],
                                                     "sensor_form": ("fused", "split", "split-tile")[best[3] - 1]},
        "vs_hit512_fp64_plain": round(v / plain, 3),
        "step_roofline": {"alg_bytes_per_cell_update": HIT_STEP_BYTES["fp32"], "achieved_GBs": round(ach, 1),
                          "frac": round(ach / HBM_PEAK_GBS, 4)},
        "form": "64 instances (16 each 128^2..512^2), FP32 state, FP64 reductions, 5 steps per HIP graph"}
    del b
    torch.cuda.empty_cache()
    return res


def fp32_error(steps, save_dir=None):
    """Production forced HIT 512x512 (seed 1234) run `steps` steps in the FP32
    batched form (FP32 state, FP64 reductions, the batch benchmark's launch
    shapes) and in FP64; relative L2 of each conserved field."""
    import dataclasses

    import numpy as np
    import torch

    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig

    c64 = HIT2DConfig.production_512(dtype="fp64", tfinal=1e9, max_steps=steps)
    c32 = dataclasses.replace(c64, dtype="fp32")
    out = {"workload": f"forced HIT 512x512 production settings, seed 1234, {steps} steps from the same initial "
                       "state; FP32 batched form vs FP64 form on the device", "steps": steps, "checkpoints": {}}
    a, b = BatchRunner([c64]), BatchRunner([c32])
    b.tune(32, 64, 2, 2)
    a.capture(10)
    b.capture(10)
    names = ("rho", "rho_u", "rho_v", "E")
    done = 0
    for ck in sorted({min(steps, x) for x in (10, 100, 500, steps)}):
        reps = (ck - done) // 10
        a.replay(reps)
        b.replay(reps)
        done += reps * 10
        torch.cuda.synchronize()
        q64, q32 = a.q[0].double(), b.q[0].double()
        err = {nm: float(torch.linalg.vector_norm(q32[k] - q64[k]) / torch.linalg.vector_norm(q64[k]))
               for k, nm in enumerate(names)}
        out["checkpoints"][str(done)] = {k: float(f"{v:.3e}") for k, v in err.items()}
    out["rel_l2_final"] = out["checkpoints"][str(done)]
    out["max_rel_l2"] = max(out["rel_l2_final"].values())
    if save_dir:
        os.makedirs(save_dir, exist_ok=True)
        np.save(os.path.join(save_dir, f"hit512_fp32_step{done}.npy"), b.q[0].cpu().numpy())
        np.save(os.path.join(save_dir, f"hit512_fp64_step{done}.npy"), a.q[0].cpu().numpy())
    return out


def serving_measurement(count, rate, steps):
    """Poisson stream of mixed requests (Riemann C3, C3 with the split at 0.8,
    C6, forced HIT; grids 128^2..1024^2; FP32/FP64) through the serving loop."""
    from hit_solver_amd.serving import ServingEngine, poisson_stream

    reqs = poisson_stream(count, rate, seed=2024, steps=tuple(steps))
    eng = ServingEngine(round_steps=10)
    res, summary = eng.run(reqs)
    cells = sum(r.n * r.n * r.steps for r in reqs)
    summary.update({
        "unit_latency": "s", "arrival_rate_per_s": rate, "seed": 2024,
        "stream": f"{count} requests, kinds uniform over c3/c3_08/c6/hit, n uniform in 128..1024 step 64, "
                  f"fp32/fp64 uniform, {steps[0]}-{steps[1]} steps each (bounded sample of the 200-5000 step "
                  f"stream), {cells / 1e9:.2f} G cell-updates in total",
        "form": "continuous batching: per-class batches on their own streams, rounds of 10 steps as HIP graphs, "
                "HBM-budget admission (80% of free device memory)"})
    for k in ("p50_s", "p99_s", "mcell_updates_per_s"):
        if summary.get(k) is not None:
            summary[k] = round(summary[k], 4)
    return summary


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "hit_solver_amd runs on exactly one GPU; multi-GPU sharding is design-only "
                                   "(instances are independent: shard the instance table)", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from hit_solver_amd.riemann import RiemannConfig3, RiemannSolver

    torch.cuda.set_device(0)
    if args.copy_calibration_mb > 0:
        src = torch.ones(args.copy_calibration_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    nx = args.nx
    cfg = RiemannConfig3(nx=nx, ny=nx, cfl=0.4, tfinal=0.3, sensor_width=4, jump_threshold=0.025,
                         shear_threshold=None, mn=0.001, hyperviscosity_interval=1, progress_every=10**9,
                         dtype=args.dtype)
    s = RiemannSolver(cfg)
    for _ in range(args.warmup):
        s.step(validate=False)
    plan = s.rhs.plan
    torch.cuda.synchronize()
    # Line-sweep kernel time: HIP events recorded around each sweep launch on
    # its stream, over eager steps that run just before the timed replays
    # (events recorded inside a captured graph cannot be read back).
    plan.timing(True)
    for _ in range(args.kernel_steps):
        s.step(validate=False)
    line_ms, line_n = plan.timing_read()
    plan.timing(False)
    s.capture()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    s.replay(args.steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_s = e0.elapsed_time(e1) / 1e3
    T = max(wall, ev_s)
    value = nx * nx * args.steps / T / 1e6
    W = 8 if args.dtype == "fp64" else 4
    # dominant kernel: the line sweeps of a stage, x and y in one launch; per sweep
    # read q (4W) + mask (1 B), write dF (4W) per cell
    line_bytes = 2 * (8 * W + 1) * nx * nx
    line_avg_s = line_ms / 1e3 / max(line_n, 1)
    achieved = line_bytes / line_avg_s / 1e9
    out = {
        "metric": "Mcell-updates/s per RK3 step",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(T / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.dtype == "fp64" else "f32",
        "data": "synthetic (Lax-Liu configuration 3 quadrant initial state)",
        "config": {"workload": f"Riemann C3 {nx}x{nx} {args.dtype} hybrid compact/WENO7, CFL 0.4, "
                               f"mn 0.001 HV every step, guard 2", "nx": nx, "ny": nx, "scheme": "hybrid",
                   "steps_full_solve": s.n_steps, "dt": s.dt},
        "roofline": {"bound": "hbm", "kernel": "k_line (x and y sweeps of a stage, one launch)",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": LINE_TRAFFIC.get((nx, args.dtype)), "launches": line_n, "avg_launch_us": round(line_avg_s * 1e6, 2),
                     "alg_bytes_per_launch": line_bytes,
                     "timing": f"HIP events around each stage's sweep launch on its stream, {args.kernel_steps} "
                               "eager steps before the timed graph replays"},
        "graph": "one step (3 fused stages + hyperviscosity) captured as a HIP graph, replayed",
        "step_roofline": {"alg_bytes_per_cell_update": STEP_BYTES[args.dtype],
                          "achieved_GBs": round(value * 1e6 * STEP_BYTES[args.dtype] / 1e9, 1),
                          "frac": round(value * 1e6 * STEP_BYTES[args.dtype] / 1e9 / HBM_PEAK_GBS, 4)},
    }
    progress(f"headline {out['value']} Mcell-updates/s")
    if not args.no_hit:
        del s
        torch.cuda.empty_cache()
        out["hit"] = hit_measurements(args.hit_steps)
        progress("forced-HIT measurements done")
    if not args.no_fp32_error:
        torch.cuda.empty_cache()
        out["fp32_error"] = fp32_error(args.fp32_steps, args.save_states)
        progress("FP32 error run done")
    if not args.no_serving:
        torch.cuda.empty_cache()
        out["serving"] = serving_measurement(args.serving_requests, args.serving_rate, args.serving_steps)
        progress("serving stream done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(nx, args.cpu_seconds)
        progress("CPU baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
