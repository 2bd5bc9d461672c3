"""Offline check of long forced-HIT runs against the oracle (test
infrastructure, run on the host): the oracle's production HIT 512x512 run
(seed 1234) for a number of steps, compared field by field with device
states saved by bench.py --save-states.

  python -m oracle.compare_hit512 --steps 1000 --oracle-cache DIR STATE.npy ...

The oracle run takes about an hour at 512x512 and 1000 steps on one core;
--oracle-cache keeps its state for reuse.
"""

from __future__ import annotations

import argparse
import os

import numpy as np

from . import hit as OH


def production_params(steps):
    return OH.HITParams(nx=512, ny=512, cfl=0.05, gamma=1.4, target_mach=0.25, initial_kmin=3, initial_kmax=5,
                        forcing_kmin=3.0, forcing_kmax=5.0, target_energy_injection=1.0e-3,
                        forcing_correlation_time=0.5, forcing_alpha_memory=0.2, min_forcing_power=1.0e-6,
                        max_forcing_rescale=20.0, mach_control=True, mach_control_target=0.25,
                        mach_control_memory=0.995, mach_control_exponent=2.0, viscosity=7.5e-4,
                        hyperviscosity_mn=0.002, large_scale_drag=0.10, large_scale_drag_kmax=2.0, cooling_time=5.0,
                        forcing_seed=1234, tfinal=110.0, max_steps=steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("states", nargs="+")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--oracle-cache", default=None)
    a = ap.parse_args()
    path = os.path.join(a.oracle_cache, f"oracle_hit512_step{a.steps}.npy") if a.oracle_cache else None
    if path and os.path.exists(path):
        ref = np.load(path)
    else:
        run = OH.HITRun(production_params(a.steps))
        run.run()
        ref = run.q
        if path:
            np.save(path, ref)
    for st in a.states:
        q = np.load(st).astype(np.float64)
        errs = [np.linalg.norm(q[k] - ref[k]) / np.linalg.norm(ref[k]) for k in range(4)]
        print(os.path.basename(st), " ".join(f"{nm}={e:.3e}" for nm, e in zip(("rho", "rho_u", "rho_v", "E"), errs)))


if __name__ == "__main__":
    main()
