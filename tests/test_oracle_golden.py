"""Oracle vs golden vectors written from the reference (tests/golden/make_golden.py).

Every comparison is bit-exact: the oracle performs the reference's NumPy
operations in the reference's order.
"""

import numpy as np
import pytest

from oracle import hit as H
from oracle import nonperiodic as N
from oracle import periodic as P
from oracle import riemann as R
from oracle.equations import Gas, enforce, viscous_rhs
from oracle.rk import ssprk3


def _eq(a, b):
    np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


def _hit_grid(nx, ny):
    p = H.HITParams(nx=nx, ny=ny)
    return p.dx, p.dy


@pytest.mark.parametrize("tag", ["hit_smooth", "hit_shock"])
def test_periodic_sensor_rhs_hv(golden, tag):
    q = golden[f"F1_{tag}_q"]
    ny, nx = q.shape[1:]
    dx, dy = _hit_grid(nx, ny)
    gas = Gas(gamma=1.4)
    sp = P.SensorParams(width=4, jump_threshold=0.04)
    m = P.sensor(enforce(q, gas), gas, dx, dy, sp)
    _eq(m, golden[f"F1_{tag}_mask"])
    _eq(P.rhs(q, gas, dx, dy, sp), golden[f"F1_{tag}_rhs"])
    assert P.stable_dt(q, gas, dx, dy, 0.05, 1.0) == golden[f"F1_{tag}_dt"]
    _eq(P.sensor(q, gas, dx, dy, P.SensorParams(width=2, jump_threshold=0.04, shear_threshold=1.5)),
        golden[f"F1_{tag}_mask_shear"])
    visc = 1.0e-3
    _eq(viscous_rhs(q, Gas(gamma=1.4, viscosity=visc), dx, dy), golden[f"F1_{tag}_viscous"])
    _eq(P.hyperviscosity(q, gas, dx, dy, 0.002, ~m), golden[f"F1_{tag}_hv"])


@pytest.mark.parametrize("tag,scheme", [("c3", "hybrid"), ("c3weno", "weno"), ("c6", "hybrid")])
def test_nonperiodic_sensor_rhs_hv(golden, tag, scheme):
    q = golden[f"F1_{tag}_q"]
    ny, nx = q.shape[1:]
    p = R.RiemannParams(nx=nx, ny=ny)
    gas = p.gas()
    sp = N.SensorParams(width=4, jump_threshold=0.025, shear_threshold=None, boundary_guard=4)
    m = N.sensor(enforce(q, gas), gas, p.dx, p.dy, sp)
    _eq(m, golden[f"F1_{tag}_mask"])
    _eq(N.rhs(q, gas, p.dx, p.dy, sp, scheme), golden[f"F1_{tag}_rhs"])
    sp2 = N.SensorParams(width=3, jump_threshold=0.025, shear_threshold=1.0, boundary_guard=2)
    _eq(N.sensor(q, gas, p.dx, p.dy, sp2), golden[f"F1_{tag}_mask_shear"])
    _eq(N.hyperviscosity(q, ~m, gas, p.dx, p.dy, 0.001), golden[f"F1_{tag}_hv"])


def test_hit_closure_and_step(golden):
    q = golden["F2_hit_q0"]
    ny, nx = q.shape[1:]
    p = H.HITParams(nx=nx, ny=ny, viscosity=7.5e-4, large_scale_drag=0.10, cooling_time=5.0)
    gas = p.gas()
    fx, fy = golden["F2_hit_fx"], golden["F2_hit_fy"]
    L = H.rhs_closure(p, gas, fx, fy)
    _eq(H.drag_source(q, gas, nx, ny, p.dx, p.dy, 0.10, 2.0), golden["F2_hit_drag"])
    _eq(H.cooling_source(q, gas, 5.0), golden["F2_hit_cooling"])
    _eq(L(q), golden["F2_hit_closure"])
    dt = P.stable_dt(q, gas, p.dx, p.dy, 0.05, 1.0)
    assert dt == golden["F2_hit_dt"]
    _eq(ssprk3(q, L, dt, clean=lambda s: enforce(s, gas)), golden["F2_hit_q1"])


def test_riemann_step_with_hv(golden):
    q = golden["F2_rie_q0"]
    ny, nx = q.shape[1:]
    p = R.config3(nx=nx, ny=ny, shear_threshold=None)
    run = R.RiemannRun(p, q0=q)
    assert run.dt == golden["F2_rie_dt"]
    q1 = ssprk3(q, run.rhs, run.dt, clean=run.clean)
    _eq(q1, golden["F2_rie_q1"])
    run.step()
    _eq(run.q, golden["F2_rie_q1_hv"])


def test_forcing_chain(golden):
    q = golden["F3_q"]
    ny, nx = q.shape[1:]
    p = H.HITParams(nx=nx, ny=ny)
    f = P  # noqa
    frc = H.ShellOUForcing(nx, ny, p.dx, p.dy, 3.0, 5.0, correlation_time=0.5, seed=1235)
    from oracle.equations import primitives

    rho, u, v, pr = primitives(q, Gas())
    keys = [str(k) for k in golden["F3_info_keys"]]
    for i, dt in enumerate([1.0e-3, 2.0e-3, 1.5e-3]):
        fx, fy, info = frc.update(dt, rho, u, v)
        _eq(fx, golden[f"F3_fx{i}"])
        _eq(fy, golden[f"F3_fy{i}"])
        _eq(np.array([info[k] for k in keys]), golden[f"F3_info{i}"])
    pm = H.HITParams(target_mach=0.25, mach_control=True, mach_control_target=0.3)
    nxt, info = H.mach_control(pm, 1.0e-3, rho, u, v, pr)
    assert nxt == golden["F3_mach_next"]
    _eq([info["mach_control_mach"], info["mach_control_target"], info["mach_control_power_desired"]],
        golden["F3_mach_info"])


def _hit_common():
    return dict(nx=32, ny=32, cfl=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5,
                forcing_kmin=3.0, forcing_kmax=5.0, forcing_seed=1234, hyperviscosity_mn=0.002)


def test_hit_run_plain(golden):
    p = H.HITParams(tfinal=0.02, viscosity=0.0, **_hit_common())
    q, t, steps = H.HITRun(p).run()
    assert steps == golden["F4_plain_steps"] and t == golden["F4_plain_t"]
    _eq(q, golden["F4_plain_q"])


def test_hit_run_production(golden):
    p = H.HITParams(tfinal=0.015, viscosity=7.5e-4, forcing_correlation_time=0.5, mach_control=True,
                    mach_control_target=0.25, large_scale_drag=0.10, cooling_time=5.0, **_hit_common())
    run = H.HITRun(p)
    alphas = []
    while not run.done():
        run.step()
        if run.steps % 5 == 0 or run.done():
            alphas.append(run.info["alpha"])
    assert run.steps == golden["F4_prod_steps"] and run.time == golden["F4_prod_t"]
    _eq(run.q, golden["F4_prod_q"])
    _eq(alphas, golden["F4_prod_hist_alpha"][1:])


@pytest.mark.parametrize("tag,factory,kw", [
    ("c3", R.config3, dict(nx=48, ny=40)),
    ("c308", R.config3_08, dict(nx=40, ny=40)),
    ("c6", R.config6, dict(nx=40, ny=48)),
])
def test_riemann_runs(golden, tag, factory, kw):
    p = factory(tfinal=0.02, shear_threshold=None, progress_every=10_000, **kw)
    run = R.RiemannRun(p)
    q = run.run()
    assert run.n_steps == golden[f"F5_{tag}_steps"]
    assert run.hv_applications == golden[f"F5_{tag}_hvapps"]
    _eq(q, golden[f"F5_{tag}_q"])
    frac = float(np.mean(N.sensor(q, run.gas, p.dx, p.dy, run.sp)))
    assert frac == golden[f"F5_{tag}_sensor_fraction"]


def test_hit_initial_state(golden):
    p = H.HITParams(nx=32, ny=24, target_mach=0.25, initial_kmin=3, initial_kmax=5)
    q = enforce(H.initial_state(p, np.random.default_rng(1234)), p.gas())
    _eq(q, golden["F6_q_32x24"])


F5_DIAG_KEYS = ("hyperviscosity_enabled", "hyperviscosity_applications", "final_sensor_fraction", "density_min",
                "density_max", "pressure_min", "pressure_max", "max_abs_vorticity", "has_nonfinite",
                "density_positive", "pressure_positive", "time", "steps")


@pytest.mark.parametrize("tag,factory,kw", [
    ("c3", R.config3, dict(nx=48, ny=40)),
    ("c308", R.config3_08, dict(nx=40, ny=40)),
    ("c6", R.config6, dict(nx=40, ny=48)),
])
def test_riemann_diagnostics_and_vorticity(golden, tag, factory, kw):
    """The oracle's run diagnostics and vorticity plane equal the reference's
    (riemann_config3.py:459-463, 629-650) on the reference's own end states."""
    p = factory(tfinal=0.02, shear_threshold=None, progress_every=10_000, **kw)
    q = golden[f"F5_{tag}_q"]
    _eq(R.vorticity_z(q, p.gas(), p.dx, p.dy), golden[f"F5_{tag}_omega"])
    steps = int(golden[f"F5_{tag}_steps"])
    d = R.diagnostics(q, p, steps * R.RiemannRun(p).dt, steps, int(golden[f"F5_{tag}_hvapps"]))
    _eq(np.array([float(d[k]) for k in F5_DIAG_KEYS]), golden[f"F5_{tag}_diag"])


def test_hit_config1_run(golden):
    """The CPU-runnable forced-HIT configuration (64x64, plain Euler, T 0.05)
    end to end, bit for bit."""
    p = H.HITParams(nx=64, ny=64, cfl=0.05, tfinal=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5,
                    forcing_kmin=3.0, forcing_kmax=5.0, forcing_seed=1234, viscosity=0.0, hyperviscosity_mn=0.002,
                    hyperviscosity_interval=5, target_energy_injection=1.0e-3, forcing_correlation_time=1.0,
                    forcing_alpha_memory=0.2)
    q, t, steps = H.HITRun(p).run()
    assert steps == golden["F7_steps"] and t == golden["F7_t"]
    _eq(q, golden["F7_q"])
