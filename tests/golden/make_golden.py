"""Write golden input/output vectors from the reference solver (HIT_Solver).

This script is the only place that imports the reference. It runs in the
build container, where the reference tree is mounted at /root/reference, and
writes small ``.npz`` fixtures next to itself.  GPU runs never need the
reference: the tests read the fixtures.

Which functions need an import.  The single-function pieces of the hot path
are pinned without the reference by closed forms and brute force in
``tests/test_oracle_pins.py`` (compact-scheme Fourier symbol, WENO7
invariants, Thomas vs dense solve, hyperviscosity symbol, viscous shear
wave, SSP-RK3 stability polynomial, initial kinetic energy, closed-form
Riemann step counts).  What only the reference can pin is the composition:

* the hybrid RHS (alpha placement in compactification, interface-mask
  convention, the transposed y treatment, clamped non-periodic gathers,
  boundary guard band)            -> F1 ``rhs_*`` / ``sensor_*``
* ``clean`` ordering inside SSP-RK3, the HIT source closure and the
  Riemann outflow guard               -> F2 ``step_*``
* the forcing chain end to end (PCG64 draw -> FFT -> shell -> OU ->
  curl -> rms -> alpha EMA/clip -> statistics)   -> F3 ``forcing_*``
* Mach control and the HIT step loop with hyperviscosity scheduling -> F4
* the Riemann fixed-dt loop with local hyperviscosity and guards    -> F5
* ``initialize_hit_2d``                                               -> F6
* the CPU-runnable forced-HIT configuration end to end                -> F7
* the Riemann run diagnostics and vorticity plane                     -> F5 ``diag`` / ``omega``

Run:  python tests/golden/make_golden.py   (numpy version is recorded)
"""

from __future__ import annotations

import contextlib
import io
import sys
import tempfile
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def _import_reference():
    sys.path.insert(0, str(REF))
    import matplotlib

    matplotlib.use("Agg")
    from OOP import domain, equations, forcing, hit2d, spatial_operator, time_operator
    from OOP.problems import riemann_config3 as rc3
    from OOP.problems import riemann_config3_08 as rc308
    from OOP.problems import riemann_config6 as rc6

    return dict(
        domain=domain,
        equations=equations,
        forcing=forcing,
        hit2d=hit2d,
        so=spatial_operator,
        time_operator=time_operator,
        rc3=rc3,
        rc308=rc308,
        rc6=rc6,
    )


def _hit_state(R, nx, ny, seed, mach, kmin=3, kmax=5, shock=False):
    """A HIT initial state (reference initializer) plus optional synthetic shock."""
    hit2d = R["hit2d"]
    cfg = hit2d.HIT2DConfig(
        nx=nx, ny=ny, target_mach=mach, initial_kmin=kmin, initial_kmax=kmax, backend="numpy"
    )
    dom, _, _ = hit2d.create_grid(cfg)
    eq = R["equations"].EulerEquation2D(gamma=1.4)
    q = hit2d.initialize_hit_2d(cfg, dom, eq, np.random.default_rng(seed))
    q = eq.enforce_physical_state(q)
    if shock:
        rho, u, v, p = eq.primitive_from_conservative(q)
        x, y = dom.mesh()
        band = (x > 2.0) & (x < 3.3) & (y > 1.0)
        rho = np.where(band, rho * 1.8, rho)
        p = np.where(band, p * 2.5, p)
        q = eq.conservative_from_primitive(rho, u, v, p)
    return cfg, dom, q


def f1_periodic(R, out):
    so = R["so"]
    eqmod = R["equations"]
    for tag, (nx, ny, shock, visc) in {
        "hit_smooth": (32, 24, False, 0.0),
        "hit_shock": (32, 24, True, 1.0e-3),
    }.items():
        cfg, dom, q = _hit_state(R, nx, ny, seed=1234, mach=0.25, shock=shock)
        eq = eqmod.EulerEquation2D(gamma=1.4)
        sensor = so.PeriodicEulerShockSensor2D(dom, eq, width=4, jump_threshold=0.04)
        op = so.PeriodicHybridEuler2DOperator(dom, eq, sensor)
        out[f"F1_{tag}_q"] = q
        out[f"F1_{tag}_mask"] = sensor.detect(eq.enforce_physical_state(q))
        out[f"F1_{tag}_rhs"] = op.rhs(q)
        out[f"F1_{tag}_dt"] = np.float64(op.stable_time_step(q, 0.05, 1.0))
        # shear branch of the sensor
        shear = so.PeriodicEulerShockSensor2D(dom, eq, width=2, jump_threshold=0.04, shear_threshold=1.5)
        out[f"F1_{tag}_mask_shear"] = shear.detect(q)
        ns = eqmod.CompressibleNavierStokes2D(gamma=1.4, viscosity=visc if visc else 1.0e-3)
        out[f"F1_{tag}_viscous"] = ns.viscous_rhs(q, dom.dx, dom.dy)
        hv = so.PeriodicHyperviscosity2D(dom, mn=0.002)
        out[f"F1_{tag}_hv"] = hv.apply(q, eq, active_mask=~out[f"F1_{tag}_mask"])


def f1_nonperiodic(R, out):
    rc3 = R["rc3"]
    for tag, (cls, nx, ny, steps, scheme) in {
        "c3": (rc3.RiemannConfig3, 56, 44, 3, "hybrid"),
        "c3weno": (rc3.RiemannConfig3, 40, 36, 2, "weno"),
        "c6": (R["rc6"].RiemannConfig6, 44, 56, 3, "hybrid"),
    }.items():
        cfg = cls(nx=nx, ny=ny, backend="numpy", scheme=scheme, shear_threshold=None)
        eq = cfg.equation
        q = rc3.apply_outflow_guard(eq.enforce_physical_state(cfg.initial_state()), cfg.guard_cells)
        sensor = rc3.NonPeriodicEulerShockSensor2D(
            cfg.domain, eq, width=4, jump_threshold=cfg.jump_threshold,
            compression_threshold=2.5, shear_threshold=None, boundary_guard=4,
        )
        op = rc3.NonPeriodicHybridEuler2DOperator(cfg.domain, eq, sensor, scheme=scheme)
        dt, _n = op.fixed_time_step(q, 0.4, 0.3)
        clean = lambda s: rc3.apply_outflow_guard(eq.enforce_physical_state(s), cfg.guard_cells)
        for _ in range(steps):  # develop some structure
            q = R["time_operator"].SSPRK3().step(q, op.rhs, dt, clean=clean)
        out[f"F1_{tag}_q"] = q
        out[f"F1_{tag}_mask"] = sensor.detect(eq.enforce_physical_state(q))
        out[f"F1_{tag}_rhs"] = op.rhs(q)
        shear = rc3.NonPeriodicEulerShockSensor2D(
            cfg.domain, eq, width=3, jump_threshold=0.025, shear_threshold=1.0, boundary_guard=2
        )
        out[f"F1_{tag}_mask_shear"] = shear.detect(q)
        hv = rc3.LocalHyperviscosity2D(cfg.domain, mn=0.001)
        out[f"F1_{tag}_hv"] = hv.apply(q, ~out[f"F1_{tag}_mask"], eq)


def f2_steps(R, out):
    hit2d = R["hit2d"]
    so = R["so"]
    cfg, dom, q = _hit_state(R, 32, 24, seed=7, mach=0.3, shock=True)
    eq = R["equations"].CompressibleNavierStokes2D(gamma=1.4, viscosity=7.5e-4, prandtl=0.72)
    sensor = so.PeriodicEulerShockSensor2D(dom, eq, width=4, jump_threshold=0.04)
    op = so.PeriodicHybridEuler2DOperator(dom, eq, sensor)
    rng = np.random.default_rng(99)
    fx = 0.1 * rng.standard_normal((dom.ny, dom.nx))
    fy = 0.1 * rng.standard_normal((dom.ny, dom.nx))

    def closure(state):
        rhs = op.rhs(state)
        rhs = rhs + eq.viscous_rhs(state, dom.dx, dom.dy)
        src = rhs + hit2d.forcing_source(state, eq, fx, fy)
        src = src + hit2d.large_scale_drag_source(state, eq, dom, 0.10, 2.0)
        src = src + hit2d.cooling_source(state, eq, 5.0, None)
        return src

    dt = op.stable_time_step(q, 0.05, 1.0)
    out["F2_hit_q0"] = q
    out["F2_hit_fx"] = fx
    out["F2_hit_fy"] = fy
    out["F2_hit_dt"] = np.float64(dt)
    out["F2_hit_closure"] = closure(q)
    out["F2_hit_drag"] = hit2d.large_scale_drag_source(q, eq, dom, 0.10, 2.0)
    out["F2_hit_cooling"] = hit2d.cooling_source(q, eq, 5.0, None)
    out["F2_hit_q1"] = R["time_operator"].SSPRK3().step(q, closure, dt, clean=eq.enforce_physical_state)

    rc3 = R["rc3"]
    cfg3 = rc3.RiemannConfig3(nx=48, ny=40, backend="numpy", shear_threshold=None)
    eq3 = cfg3.equation
    q3 = rc3.apply_outflow_guard(eq3.enforce_physical_state(cfg3.initial_state()), 2)
    sensor3 = rc3.NonPeriodicEulerShockSensor2D(cfg3.domain, eq3, width=4, jump_threshold=0.025,
                                               shear_threshold=None, boundary_guard=4)
    op3 = rc3.NonPeriodicHybridEuler2DOperator(cfg3.domain, eq3, sensor3)
    dt3, _ = op3.fixed_time_step(q3, 0.4, 0.3)
    clean = lambda s: rc3.apply_outflow_guard(eq3.enforce_physical_state(s), 2)
    q3a = R["time_operator"].SSPRK3().step(q3, op3.rhs, dt3, clean=clean)
    out["F2_rie_q0"] = q3
    out["F2_rie_dt"] = np.float64(dt3)
    out["F2_rie_q1"] = q3a
    hv = rc3.LocalHyperviscosity2D(cfg3.domain, mn=0.001)
    out["F2_rie_q1_hv"] = rc3.apply_outflow_guard(hv.apply(q3a, ~sensor3.detect(q3a), eq3), 2)


def f3_forcing(R, out):
    hit2d = R["hit2d"]
    cfg, dom, q = _hit_state(R, 32, 24, seed=1234, mach=0.25)
    eq = R["equations"].EulerEquation2D(gamma=1.4)
    frc = R["forcing"].IsotropicShellOUForcing2D(
        domain=dom, k_min=3.0, k_max=5.0, correlation_time=0.5, force_rms=1.0,
        target_power=1.0e-3, min_power=1.0e-6, max_rescale=20.0, alpha_memory=0.2, seed=1235,
    )
    rho, u, v, _p = eq.primitive_from_conservative(q)
    out["F3_q"] = q
    keys = ["alpha", "alpha_target", "power_before_rescale", "injected_power", "Fxx", "Fyy", "Fxy", "A_F", "ou_decay"]
    for i, dt in enumerate([1.0e-3, 2.0e-3, 1.5e-3]):
        fx, fy, info = frc.update(dt, rho, u, v)
        out[f"F3_fx{i}"] = fx
        out[f"F3_fy{i}"] = fy
        out[f"F3_info{i}"] = np.array([info[k] for k in keys])
    out["F3_info_keys"] = np.array(keys)
    # Mach control, one update from this state
    cfgm = hit2d.HIT2DConfig(target_mach=0.25, mach_control=True, mach_control_target=0.3)
    rho, u, v, p = eq.primitive_from_conservative(q)
    nxt, info = hit2d.update_mach_controlled_power(cfgm, 1.0e-3, rho, u, v, p)
    out["F3_mach_next"] = np.float64(nxt)
    out["F3_mach_info"] = np.array([info["mach_control_mach"], info["mach_control_target"],
                                    info["mach_control_power_desired"]])


def _run_hit(R, tmp, **kw):
    hit2d = R["hit2d"]
    cfg = hit2d.HIT2DConfig(output_dir=Path(tmp), backend="numpy", **kw)
    with contextlib.redirect_stdout(io.StringIO()):
        q, t, steps = hit2d.run_simulation(cfg)
    hist = np.load(Path(tmp) / "diagnostic_history.npz", allow_pickle=False)
    return q, t, steps, hist


def f4_hit_runs(R, out):
    common = dict(
        nx=32, ny=32, cfl=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5,
        forcing_kmin=3.0, forcing_kmax=5.0, forcing_seed=1234, hyperviscosity_mn=0.002,
        diagnostics_every=5, snapshot_every=10_000,
    )
    with tempfile.TemporaryDirectory() as tmp:
        # the small CPU-runnable HIT case: plain Euler, no drag/cooling, tfinal 0.02 (~12 steps, 2 HV passes)
        q, t, steps, hist = _run_hit(R, tmp, tfinal=0.02, viscosity=0.0, **common)
        out["F4_plain_q"] = q
        out["F4_plain_t"] = np.float64(t)
        out["F4_plain_steps"] = np.int64(steps)
        out["F4_plain_hist_KE"] = hist["kinetic_energy"]
    with tempfile.TemporaryDirectory() as tmp:
        # production settings (viscosity, Mach control, drag, cooling), short
        q, t, steps, hist = _run_hit(
            R, tmp, tfinal=0.015, viscosity=7.5e-4, forcing_correlation_time=0.5,
            mach_control=True, mach_control_target=0.25, large_scale_drag=0.10,
            cooling_time=5.0, **common,
        )
        out["F4_prod_q"] = q
        out["F4_prod_t"] = np.float64(t)
        out["F4_prod_steps"] = np.int64(steps)
        out["F4_prod_hist_Pin"] = hist["P_in"]
        out["F4_prod_hist_alpha"] = hist["forcing_alpha"]


F5_DIAG_KEYS = ("hyperviscosity_enabled", "hyperviscosity_applications", "final_sensor_fraction", "density_min",
                "density_max", "pressure_min", "pressure_max", "max_abs_vorticity", "has_nonfinite",
                "density_positive", "pressure_positive", "time", "steps")


def f5_riemann_runs(R, out):
    rc3 = R["rc3"]
    for tag, (cls, kw) in {
        "c3": (rc3.RiemannConfig3, dict(nx=48, ny=40, tfinal=0.02, shear_threshold=None)),
        "c308": (R["rc308"].RiemannConfig3Offset08, dict(nx=40, ny=40, tfinal=0.02, shear_threshold=None)),
        "c6": (R["rc6"].RiemannConfig6, dict(nx=40, ny=48, tfinal=0.02, shear_threshold=None)),
    }.items():
        cfg = cls(backend="numpy", progress_every=10_000, **kw)
        with contextlib.redirect_stdout(io.StringIO()):
            q, diag = rc3.run_riemann(cfg)
        out[f"F5_{tag}_q"] = q
        out[f"F5_{tag}_steps"] = np.int64(diag.steps)
        out[f"F5_{tag}_hvapps"] = np.int64(diag.hyperviscosity_applications)
        out[f"F5_{tag}_sensor_fraction"] = np.float64(diag.final_sensor_fraction)
        # the reference's end-of-run diagnostics (riemann_config3.py:629-650)
        # and its vorticity plane (one-sided differences at the ends, :459-463)
        out[f"F5_{tag}_diag"] = np.array([float(getattr(diag, k)) for k in F5_DIAG_KEYS])
        out[f"F5_{tag}_omega"] = rc3.vorticity_z(q, cfg.domain, cfg.equation)


def f7_config1(R, out):
    """The CPU-runnable forced-HIT case as the benchmark states it: 64x64,
    FP64, plain Euler (viscosity 0), Mach 0.25, shells 3-5, seed 1234, CFL
    0.05, T 0.05, hyperviscosity mn 0.002 every 5 steps, P_target 1e-3,
    tau_OU 1.0, alpha memory 0.2, no drag, cooling or Mach control."""
    with tempfile.TemporaryDirectory() as tmp:
        q, t, steps, hist = _run_hit(
            R, tmp, nx=64, ny=64, cfl=0.05, tfinal=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5,
            forcing_kmin=3.0, forcing_kmax=5.0, forcing_seed=1234, viscosity=0.0, hyperviscosity_mn=0.002,
            hyperviscosity_interval=5, target_energy_injection=1.0e-3, forcing_correlation_time=1.0,
            forcing_alpha_memory=0.2, diagnostics_every=10, snapshot_every=10_000,
        )
    out["F7_q"] = q
    out["F7_t"] = np.float64(t)
    out["F7_steps"] = np.int64(steps)
    out["F7_hist_KE"] = hist["kinetic_energy"]


def f6_init(R, out):
    hit2d = R["hit2d"]
    cfg, dom, q = _hit_state(R, 32, 24, seed=1234, mach=0.25)
    out["F6_q_32x24"] = q
    # 512^2 initial scalars, compared with the reference's own run.log line 2
    cfg = hit2d.HIT2DConfig(nx=512, ny=512, target_mach=0.25, initial_kmin=3, initial_kmax=5)
    dom, _, _ = hit2d.create_grid(cfg)
    eq = R["equations"].EulerEquation2D(gamma=1.4)
    q = eq.enforce_physical_state(hit2d.initialize_hit_2d(cfg, dom, eq, np.random.default_rng(1234)))
    sensor = R["so"].PeriodicEulerShockSensor2D(dom, eq, width=4, jump_threshold=0.04)
    d = hit2d.compute_diagnostics(q, eq, dom, weno_fraction=float(np.mean(sensor.detect(q))))
    out["F6_512_scalars"] = np.array([d["kinetic_energy"], d["turbulent_mach"], d["weno_fraction"],
                                      d["vorticity_rms"], d["divergence_rms"]])
    out["F6_512_mask_count"] = np.int64(sensor.detect(q).sum())


def main():
    R = _import_reference()
    out: dict[str, np.ndarray] = {"numpy_version": np.array(np.__version__)}
    f1_periodic(R, out)
    f1_nonperiodic(R, out)
    f2_steps(R, out)
    f3_forcing(R, out)
    f4_hit_runs(R, out)
    f5_riemann_runs(R, out)
    f6_init(R, out)
    f7_config1(R, out)
    path = OUT / "reference_vectors.npz"
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({path.stat().st_size/1024:.1f} KiB, {len(out)} arrays)")


if __name__ == "__main__":
    main()
