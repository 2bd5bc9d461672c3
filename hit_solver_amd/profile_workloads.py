"""Fixed workloads for profiling under rocprofv3 (no timing of its own).

  python3 -m hit_solver_amd.profile_workloads batch --steps 8 --tiles 32 64 2 2
  python3 -m hit_solver_amd.profile_workloads riemann --steps 20 [--n 1024]
  python3 -m hit_solver_amd.profile_workloads hit --steps 10

batch: 64 forced-HIT instances, 16 each at 128^2, 256^2, 384^2, 512^2, seeds
0-63, Mach (0.1, 0.25, 0.4)[seed % 3], production settings, FP32, with the
given launch shapes (rb, rbp, line form, sensor form) instead of the
autotuner, so every profiled step uses the same kernels.
riemann: Riemann configuration 3 at 512^2 (or --n) FP64, the single-plan driver.
hit: production forced HIT 512^2 FP64, one instance on device control.
"""

from __future__ import annotations

import argparse

import torch


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("batch", "riemann", "hit"))
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dtype", default=None)
    ap.add_argument("--n", type=int, default=512, help="grid size of the riemann workload")
    ap.add_argument("--tiles", type=int, nargs=4, default=(32, 64, 2, 2), metavar=("RB", "RBP", "LINE", "SENSOR"))
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="3 plain device copies of this many MiB first (scale for FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args(argv)
    torch.cuda.set_device(0)
    if a.copy_calibration_mb > 0:
        src = torch.ones(a.copy_calibration_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if a.workload == "batch":
        from .batch import BatchRunner, hit_batched_configs

        b = BatchRunner(hit_batched_configs(a.dtype or "fp32"))
        b.tune(*a.tiles)
        step = b.step
    elif a.workload == "riemann":
        from .riemann import RiemannConfig3, RiemannSolver

        s = RiemannSolver(RiemannConfig3(nx=a.n, ny=a.n, shear_threshold=None, progress_every=10**9,
                                         dtype=a.dtype or "fp64"))
        step = lambda: s.step(validate=False)  # noqa: E731
    else:
        from .batch import BatchRunner
        from .hit2d import HIT2DConfig

        b = BatchRunner([HIT2DConfig.production_512(dtype=a.dtype or "fp64", tfinal=1e9)])
        step = b.step
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    print(f"{a.workload}: {a.warmup} warm-up + {a.steps} profiled steps", flush=True)


if __name__ == "__main__":
    main()
