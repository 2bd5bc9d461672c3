"""The benchmark configurations as written, on the device, against the oracle.

* Config 1: forced HIT 64x64, FP64, plain Euler (no viscosity), Mach 0.25,
  streamfunction shell 3-5 (seed 1234), OU forcing in 3-5 (noise seed 1235),
  CFL 0.05, T 0.05, hyperviscosity mn 0.002 every 5 steps, P_target 1e-3,
  tau_OU 1.0, memory 0.2, no drag, cooling or Mach control; against the
  NumPy FP64 oracle run of the same configuration (relative L2 1e-10 per
  field: a smooth flow, no WENO amplification).
* FP32 with FP64 reductions against the FP64 oracle: production forced HIT
  (viscosity 7.5e-4, OU tau 0.5, Mach control, drag 0.10 for k <= 2, cooling
  5.0, hyperviscosity every 5 steps), as one instance (the Python operator
  API) and as a three-instance FP32 batch of uneven grids.  The FP32 error
  grows about linearly with the step count (forced HIT 512x512: 4.7e-7 at 10
  steps, 5.1e-6 at 100, 6.3e-5 at 1000); after 50 steps the per-field bound
  is 1e-5.
* Config 4 at full size: the 64-instance FP32 batch (16 each at 128^2,
  256^2, 384^2, 512^2; seeds 0-63; Mach (0.1, 0.25, 0.4)[seed % 3]) for 20
  steps, against the same 64 instances run in FP64 on the device; per-field
  relative L2 <= 5e-6 for every instance (2e-6 expected at 20 steps from the
  growth above).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


CONFIG1 = dict(nx=64, ny=64, cfl=0.05, tfinal=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5,
               forcing_kmin=3.0, forcing_kmax=5.0, forcing_seed=1234, viscosity=0.0, hyperviscosity_mn=0.002,
               hyperviscosity_interval=5, target_energy_injection=1.0e-3, forcing_correlation_time=1.0,
               forcing_alpha_memory=0.2)


def production(**kw):
    base = dict(cfl=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0,
                target_energy_injection=1.0e-3, forcing_correlation_time=0.5, forcing_alpha_memory=0.2,
                mach_control=True, mach_control_target=0.25, mach_control_memory=0.995, viscosity=7.5e-4,
                hyperviscosity_mn=0.002, hyperviscosity_interval=5, large_scale_drag=0.10, large_scale_drag_kmax=2.0,
                cooling_time=5.0, tfinal=1.0)
    base.update(kw)
    return base


def test_config1_hit64_vs_oracle():
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver
    from oracle import hit as OH

    s = HITSolver(HIT2DConfig(**CONFIG1))
    s.run()
    ref = OH.HITRun(OH.HITParams(**CONFIG1))
    ref.run()
    print(f"[configs] config 1: {s.steps} steps, t {s.time:.6f}", flush=True)
    assert s.steps == ref.steps and s.steps >= 25
    assert s.time == pytest.approx(ref.time, rel=1e-13)
    q = host(s.q)
    for k in range(4):
        assert rel_l2(q[k], ref.q[k]) <= 1e-10, k


def test_fp32_production_single_vs_oracle():
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver
    from oracle import hit as OH

    kw = production(nx=64, ny=64, forcing_seed=21, max_steps=50)
    s = HITSolver(HIT2DConfig(dtype="fp32", **kw))
    s.run()
    ref = OH.HITRun(OH.HITParams(**kw))
    ref.run()
    assert s.steps == ref.steps == 50
    q = host(s.q)
    errs = [rel_l2(q[k], ref.q[k]) for k in range(4)]
    print(f"[configs] FP32 single 64^2 x50 vs oracle: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-5


def test_fp32_production_batch_vs_oracle():
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig
    from oracle import hit as OH

    kws = [production(nx=48, ny=48, forcing_seed=31, target_mach=0.1, mach_control_target=0.1, max_steps=50),
           production(nx=64, ny=96, forcing_seed=32, max_steps=50),
           production(nx=96, ny=80, forcing_seed=33, target_mach=0.4, mach_control_target=0.4, max_steps=50)]
    b = BatchRunner([HIT2DConfig(dtype="fp32", **kw) for kw in kws])
    b.run()
    ctl = b.control()
    for i, kw in enumerate(kws):
        ref = OH.HITRun(OH.HITParams(**kw))
        ref.run()
        assert int(ctl[i, 4]) == ref.steps == 50
        q = host(b.q[i])
        errs = [rel_l2(q[k], ref.q[k]) for k in range(4)]
        print(f"[configs] FP32 batch instance {i} ({kw['nx']}x{kw['ny']}) x50 vs oracle: "
              f"{['%.2e' % e for e in errs]}", flush=True)
        assert max(errs) <= 1e-5, (i, errs)


def test_config4_full_batch_fp32_vs_fp64():
    from hit_solver_amd.batch import BatchRunner, hit_batched_configs

    steps = 20
    c32, c64 = hit_batched_configs("fp32"), hit_batched_configs("fp64")
    assert len(c32) == 64 and sum(c.nx * c.ny for c in c32) == 7_864_320
    print("[configs] config 4: 64 instances FP64, 20 steps", flush=True)
    b64 = BatchRunner(c64, max_steps=steps)
    b64.run()
    q64 = [host(q) for q in b64.q]
    ctl64 = b64.control()
    del b64
    torch.cuda.empty_cache()
    print("[configs] config 4: 64 instances FP32 (autotuned), 20 steps", flush=True)
    b32 = BatchRunner(c32, max_steps=steps)
    b32.autotune()
    b32.run()
    ctl32 = b32.control()
    worst = {}
    for i in range(64):
        assert ctl32[i, 4] == ctl64[i, 4] == steps
        q = host(b32.q[i])
        errs = [rel_l2(q[k], q64[i][k]) for k in range(4)]
        n = c32[i].nx
        worst[n] = max(worst.get(n, 0.0), max(errs))
        assert max(errs) <= 2e-5, (i, n, errs)
    print(f"[configs] config 4 worst per-field rel L2 FP32 vs FP64 after {steps} steps, by grid: "
          f"{ {n: f'{e:.2e}' for n, e in sorted(worst.items())} }", flush=True)
