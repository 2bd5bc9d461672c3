"""HIP path vs the reference's golden vectors and the NumPy oracle.

Tolerances (FP64):
* shock-sensor masks, hyperviscosity, viscous terms, enforce: bit-exact
  (same arithmetic in the same order, no FMA contraction);
* operator RHS: max |diff| <= 1e-12 * max |rhs| (the reference solves the
  periodic compact system by FFT and the non-periodic one by sequential
  Thomas; the kernels use a chunked recurrence -- equal to rounding);
* one SSP-RK3 step: relative L2 <= 1e-12 per conserved field;
* multi-step runs: relative L2 <= 10x the oracle's own divergence under a
  1-ulp perturbation of the initial state, measured in the same test (floor
  1e-12).  WENO7's smoothness indicators amplify rounding noise on nearly
  flat data (the reference's CPU and GPU runs diverge the same way): on
  Riemann C3 at 48x40 the oracle moves 1.7e-10 from itself in 8 steps.
"""

import numpy as np
import pytest
import torch

from oracle import hit as OH
from oracle import nonperiodic as ON
from oracle import periodic as OP
from oracle import riemann as OR
from oracle.equations import Gas, enforce as o_enforce, viscous_rhs as o_viscous

pytestmark = pytest.mark.gpu

DEV = "cuda"


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def host(t):
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def assert_fields(a, b, tol, what=""):
    for k in range(4):
        e = rel_l2(a[k], b[k])
        assert e <= tol, f"{what} field {k}: rel L2 {e:.3e} > {tol:.1e}"


def ulp_sensitivity(make_run, steps, seed=0):
    """Per-field rel-L2 between two oracle runs whose initial states differ
    by one ulp per entry."""
    a, b = make_run(), make_run()
    rng = np.random.default_rng(seed)
    up = rng.random(b.q.shape) < 0.5
    b.q = np.where(up, np.nextafter(b.q, np.inf), np.nextafter(b.q, -np.inf))
    for r in (a, b):
        if hasattr(r, "n_steps"):
            r.run(max_steps=steps)
        else:
            while not r.done():
                r.step()
    return [rel_l2(b.q[k], a.q[k]) for k in range(4)], a


def assert_within_sensitivity(got, sens, ref_q, what, factor=10.0):
    for k in range(4):
        e = rel_l2(got[k], ref_q[k])
        lim = max(factor * sens[k], 1e-12)
        assert e <= lim, f"{what} field {k}: rel L2 {e:.3e} > {factor:g} x 1-ulp sensitivity {sens[k]:.3e}"


def assert_rhs(a, b, tol=1e-12, what=""):
    for k in range(4):
        scale = max(np.max(np.abs(b[k])), 1e-300)
        e = np.max(np.abs(a[k] - b[k])) / scale
        assert e <= tol, f"{what} component {k}: max|diff|/max|ref| {e:.3e} > {tol:.1e}"


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from hit_solver_amd._lib import lib

    lib()


def _hit_objs(nx, ny, viscosity=0.0, width=4, shear=None):
    from hit_solver_amd.domain import Domain2D
    from hit_solver_amd.equations import CompressibleNavierStokes2D, EulerEquation2D
    from hit_solver_amd.spatial_operator import (PeriodicEulerShockSensor2D, PeriodicHybridEuler2DOperator,
                                                 PeriodicHyperviscosity2D)

    d = Domain2D(0.0, 2 * np.pi, 0.0, 2 * np.pi, nx, ny)
    eq = CompressibleNavierStokes2D(gamma=1.4, viscosity=viscosity) if viscosity else EulerEquation2D(gamma=1.4)
    s = PeriodicEulerShockSensor2D(d, eq, width=width, jump_threshold=0.04, shear_threshold=shear)
    return d, eq, s, PeriodicHybridEuler2DOperator(d, eq, s), PeriodicHyperviscosity2D(d, mn=0.002)


def _rie_objs(nx, ny, scheme="hybrid", width=4, shear=None, bguard=4):
    from hit_solver_amd.riemann import (LocalHyperviscosity2D, NonPeriodicEulerShockSensor2D,
                                        NonPeriodicHybridEuler2DOperator, RiemannConfig3)

    cfg = RiemannConfig3(nx=nx, ny=ny, scheme=scheme)
    d, eq = cfg.domain, cfg.equation
    s = NonPeriodicEulerShockSensor2D(d, eq, width=width, jump_threshold=0.025, shear_threshold=shear,
                                      boundary_guard=bguard)
    return d, eq, s, NonPeriodicHybridEuler2DOperator(d, eq, s, scheme=scheme), LocalHyperviscosity2D(d, mn=0.001)


# ---------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("tag", ["hit_smooth", "hit_shock"])
def test_periodic_sensor_rhs_hv_golden(golden, tag):
    q = golden[f"F1_{tag}_q"]
    ny, nx = q.shape[1:]
    d, eq, s, op, hv = _hit_objs(nx, ny)
    qd = dev(q)
    m = host(s.detect(eq.enforce_physical_state(qd)))
    np.testing.assert_array_equal(m, golden[f"F1_{tag}_mask"])
    assert_rhs(host(op.rhs(qd)), golden[f"F1_{tag}_rhs"], what=tag)
    assert op.stable_time_step(qd, 0.05, 1.0) == pytest.approx(float(golden[f"F1_{tag}_dt"]), rel=1e-15)
    _, _, s2, _, _ = _hit_objs(nx, ny, width=2, shear=1.5)
    np.testing.assert_array_equal(host(s2.detect(qd)), golden[f"F1_{tag}_mask_shear"])
    _, eqv, _, _, _ = _hit_objs(nx, ny, viscosity=1.0e-3)
    np.testing.assert_array_equal(host(eqv.viscous_rhs(qd, d.dx, d.dy)), golden[f"F1_{tag}_viscous"])
    act = dev(~golden[f"F1_{tag}_mask"], torch.uint8)
    np.testing.assert_array_equal(host(hv.apply(qd, eq, active_mask=act)), golden[f"F1_{tag}_hv"])


@pytest.mark.parametrize("tag,scheme", [("c3", "hybrid"), ("c3weno", "weno"), ("c6", "hybrid")])
def test_nonperiodic_sensor_rhs_hv_golden(golden, tag, scheme):
    q = golden[f"F1_{tag}_q"]
    ny, nx = q.shape[1:]
    d, eq, s, op, hv = _rie_objs(nx, ny, scheme=scheme)
    qd = dev(q)
    np.testing.assert_array_equal(host(s.detect(eq.enforce_physical_state(qd))), golden[f"F1_{tag}_mask"])
    assert_rhs(host(op.rhs(qd)), golden[f"F1_{tag}_rhs"], what=tag)
    _, _, s2, _, _ = _rie_objs(nx, ny, width=3, shear=1.0, bguard=2)
    np.testing.assert_array_equal(host(s2.detect(qd)), golden[f"F1_{tag}_mask_shear"])
    act = dev(~golden[f"F1_{tag}_mask"], torch.uint8)
    np.testing.assert_array_equal(host(hv.apply(qd, act, eq)), golden[f"F1_{tag}_hv"])


def test_enforce_bitexact(golden):
    from hit_solver_amd.equations import EulerEquation2D

    q = golden["F1_hit_shock_q"].copy()
    q[0, 3, 4] = -1.0  # floor density
    q[3, 5, 6] = 0.0   # floor pressure
    got = host(EulerEquation2D().enforce_physical_state(dev(q)))
    np.testing.assert_array_equal(got, o_enforce(q, Gas()))


def test_hit_closure_step_golden(golden):
    from hit_solver_amd.hit2d import drag_modes
    from hit_solver_amd.time_operator import NativeRHS, SSPRK3

    q = golden["F2_hit_q0"]
    ny, nx = q.shape[1:]
    d, eq, s, op, _ = _hit_objs(nx, ny, viscosity=7.5e-4)
    dm = drag_modes(d, 2.0)
    plan = op.plan(torch.float64, viscosity=7.5e-4, forcing=True, drag_coef=0.10,
                   drag_modes=tuple(int(v) for v in dm.reshape(-1)), cooling_time=5.0)
    rhs = NativeRHS(plan, dev(golden["F2_hit_fx"]), dev(golden["F2_hit_fy"]))
    qd = dev(q)
    dt = op.stable_time_step(qd, 0.05, 1.0)
    assert dt == pytest.approx(float(golden["F2_hit_dt"]), rel=1e-15)
    out = host(SSPRK3().step(qd, rhs, float(golden["F2_hit_dt"]), clean=rhs.clean))
    assert_fields(out, golden["F2_hit_q1"], 1e-12, "HIT step")


def test_riemann_step_hv_golden(golden):
    from hit_solver_amd.riemann import RiemannConfig3, RiemannSolver

    q = golden["F2_rie_q0"]
    ny, nx = q.shape[1:]
    solver = RiemannSolver(RiemannConfig3(nx=nx, ny=ny, shear_threshold=None), q0=dev(q))
    assert solver.dt == pytest.approx(float(golden["F2_rie_dt"]), rel=1e-15)
    solver.dt = float(golden["F2_rie_dt"])
    solver.step()
    assert_fields(host(solver.q), golden["F2_rie_q1_hv"], 1e-12, "Riemann step + HV")


def test_forcing_chain_golden(golden):
    from hit_solver_amd.domain import Domain2D
    from hit_solver_amd.equations import EulerEquation2D
    from hit_solver_amd.forcing import IsotropicShellOUForcing2D

    q = golden["F3_q"]
    ny, nx = q.shape[1:]
    d = Domain2D(0.0, 2 * np.pi, 0.0, 2 * np.pi, nx, ny)
    f = IsotropicShellOUForcing2D(domain=d, k_min=3.0, k_max=5.0, correlation_time=0.5, seed=1235)
    rho, u, v, _ = EulerEquation2D().primitive_from_conservative(dev(q))
    keys = [str(k) for k in golden["F3_info_keys"]]
    for i, dt in enumerate([1.0e-3, 2.0e-3, 1.5e-3]):
        fx, fy, info = f.update(dt, rho, u, v)
        np.testing.assert_allclose(host(fx), golden[f"F3_fx{i}"], rtol=0, atol=1e-12 * np.abs(golden[f"F3_fx{i}"]).max())
        np.testing.assert_allclose(host(fy), golden[f"F3_fy{i}"], rtol=0, atol=1e-12 * np.abs(golden[f"F3_fy{i}"]).max())
        np.testing.assert_allclose([info[k] for k in keys], golden[f"F3_info{i}"], rtol=1e-11, atol=1e-15)


def _hit_common():
    return dict(nx=32, ny=32, cfl=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0,
                forcing_kmax=5.0, forcing_seed=1234, hyperviscosity_mn=0.002)


def test_hit_run_plain_golden(golden):
    from hit_solver_amd.hit2d import HIT2DConfig, run_simulation

    q, t, steps = run_simulation(HIT2DConfig(tfinal=0.02, viscosity=0.0, **_hit_common()))
    assert steps == int(golden["F4_plain_steps"])
    assert t == pytest.approx(float(golden["F4_plain_t"]), rel=1e-13)
    assert_fields(host(q), golden["F4_plain_q"], 1e-10, "HIT plain run")  # smooth flow: no amplification


def test_hit_run_production_golden(golden):
    from hit_solver_amd.hit2d import HIT2DConfig, run_simulation

    cfg = HIT2DConfig(tfinal=0.015, viscosity=7.5e-4, forcing_correlation_time=0.5, mach_control=True,
                      mach_control_target=0.25, large_scale_drag=0.10, cooling_time=5.0, **_hit_common())
    q, t, steps = run_simulation(cfg)
    assert steps == int(golden["F4_prod_steps"])
    assert_fields(host(q), golden["F4_prod_q"], 1e-10, "HIT production run")


@pytest.mark.parametrize("tag,cls,kw", [
    ("c3", "RiemannConfig3", dict(nx=48, ny=40)),
    ("c308", "RiemannConfig3Offset08", dict(nx=40, ny=40)),
    ("c6", "RiemannConfig6", dict(nx=40, ny=48)),
])
def test_riemann_runs_golden(golden, tag, cls, kw):
    import hit_solver_amd.riemann as R

    cfg = getattr(R, cls)(tfinal=0.02, shear_threshold=None, progress_every=10_000, **kw)
    q, diag = R.run_riemann(cfg)
    assert diag.steps == int(golden[f"F5_{tag}_steps"])
    assert diag.hyperviscosity_applications == int(golden[f"F5_{tag}_hvapps"])
    factory = {"c3": OR.config3, "c308": OR.config3_08, "c6": OR.config6}[tag]
    p = factory(tfinal=0.02, shear_threshold=None, progress_every=10_000, **kw)
    sens, ref = ulp_sensitivity(lambda: OR.RiemannRun(p), diag.steps)
    np.testing.assert_array_equal(ref.q, golden[f"F5_{tag}_q"])  # the oracle run is the reference's
    assert_within_sensitivity(host(q), sens, golden[f"F5_{tag}_q"], f"Riemann {tag}")
    assert abs(diag.final_sensor_fraction - float(golden[f"F5_{tag}_sensor_fraction"])) <= 2.0 / (cfg.nx * cfg.ny)


# ---------------------------------------------------------------- vs oracle, other sizes
@pytest.mark.parametrize("nx,ny", [(64, 48), (36, 42), (37, 40), (128, 96), (192, 128)])
def test_periodic_rhs_vs_oracle_sizes(nx, ny):
    p = OH.HITParams(nx=nx, ny=ny, target_mach=0.3, initial_kmin=3, initial_kmax=5)
    q = OH.initial_state(p, np.random.default_rng(nx * 1000 + ny))
    x = np.arange(nx) * p.dx
    q[0] *= 1.0 + 0.6 * (np.sin(x)[None, :] > 0.3)  # density steps -> WENO bands
    q = o_enforce(q, Gas())
    d, eq, s, op, _ = _hit_objs(nx, ny)
    qd = dev(q)
    sp = OP.SensorParams()
    np.testing.assert_array_equal(host(s.detect(eq.enforce_physical_state(qd))),
                                  OP.sensor(o_enforce(q, Gas()), Gas(), p.dx, p.dy, sp))
    assert_rhs(host(op.rhs(qd)), OP.rhs(q, Gas(), p.dx, p.dy, sp), what=f"{nx}x{ny}")


@pytest.mark.parametrize("nx,ny,scheme", [(64, 64, "hybrid"), (50, 36, "hybrid"), (33, 40, "weno"),
                                          (128, 192, "hybrid"), (64, 128, "weno")])
def test_nonperiodic_rhs_vs_oracle_sizes(nx, ny, scheme):
    p = OR.config3(nx=nx, ny=ny, shear_threshold=None, scheme=scheme)
    run = OR.RiemannRun(p)
    run.run(max_steps=3)
    q = run.q
    d, eq, s, op, _ = _rie_objs(nx, ny, scheme=scheme)
    sp = p.sensor_params()
    np.testing.assert_array_equal(host(s.detect(dev(q))), ON.sensor(q, Gas(), p.dx, p.dy, sp))
    assert_rhs(host(op.rhs(dev(q))), ON.rhs(q, Gas(), p.dx, p.dy, sp, scheme), what=f"{nx}x{ny}")


@pytest.mark.parametrize("case", ["periodic", "riemann"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sensor_forms_agree(case, dtype):
    """The fused tile sensor and the split form (raw-mask column walk, then
    the dilation tile) give the same mask, bit for bit, and the same rhs."""
    if case == "periodic":
        nx, ny = 96, 80
        p = OH.HITParams(nx=nx, ny=ny, target_mach=0.3, initial_kmin=3, initial_kmax=5)
        q = OH.initial_state(p, np.random.default_rng(3))
        x = np.arange(nx) * p.dx
        q[0] *= 1.0 + 0.6 * (np.sin(x)[None, :] > 0.3)
        q = o_enforce(q, Gas())
        _, _, sensor, op, _ = _hit_objs(nx, ny)
    else:
        nx, ny = 72, 56
        p = OR.config3(nx=nx, ny=ny, shear_threshold=None)
        run = OR.RiemannRun(p)
        run.run(max_steps=3)
        q = run.q
        _, _, sensor, op, _ = _rie_objs(nx, ny)
    qd = dev(q, dtype)
    plan = op.plan(dtype)
    out = {}
    for form in (1, 2):
        plan.set_sensor_form(form)
        out[form] = (host(plan.sensor(qd, enforce_first=True)), host(op.rhs(qd)))
    plan.set_sensor_form(0)
    np.testing.assert_array_equal(out[1][0], out[2][0])
    assert out[1][0].any() and not out[1][0].all()
    np.testing.assert_array_equal(out[1][1], out[2][1])


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-6)])
@pytest.mark.parametrize("case", ["periodic", "nonperiodic"])
def test_line_forms_agree(case, dtype, tol):
    """The one-wavefront-per-line sweep and the block sweep give the same
    derivative (rounding only: the wave form's solve is an exact scan, the
    block form's truncates |rho|^52) on lines of 64*L nodes, L = 1..8."""
    if case == "periodic":
        nx, ny = 512, 320
        p = OH.HITParams(nx=nx, ny=ny, target_mach=0.3, initial_kmin=3, initial_kmax=5)
        q = OH.initial_state(p, np.random.default_rng(7))
        x = np.arange(nx) * p.dx
        q[0] *= 1.0 + 0.6 * (np.sin(x)[None, :] > 0.3)
        q = o_enforce(q, Gas())
        _, _, _, op, _ = _hit_objs(nx, ny)
    else:
        nx, ny = 448, 64
        p = OR.config3(nx=nx, ny=ny, shear_threshold=None)
        run = OR.RiemannRun(p)
        run.run(max_steps=2)
        q = run.q
        _, _, _, op, _ = _rie_objs(nx, ny)
    qd = dev(q, dtype)
    plan = op.plan(dtype)
    plan.set_line_form(1)
    a = host(op.rhs(qd)).astype(np.float64)
    plan.set_line_form(2)
    b = host(op.rhs(qd)).astype(np.float64)
    plan.set_line_form(0)
    for c in range(4):
        scale = np.abs(a[c]).max()
        assert np.abs(b[c] - a[c]).max() <= tol * scale, (case, c, np.abs(b[c] - a[c]).max() / scale)


def test_riemann_oracle_multistep_128():
    from hit_solver_amd.riemann import RiemannConfig3, RiemannSolver

    p = OR.config3(nx=128, ny=128, shear_threshold=None, progress_every=10_000)
    sens, ref = ulp_sensitivity(lambda: OR.RiemannRun(p), 10)
    s = RiemannSolver(RiemannConfig3(nx=128, ny=128, shear_threshold=None, progress_every=10_000))
    assert s.n_steps == ref.n_steps and s.dt == pytest.approx(ref.dt, rel=1e-15)
    s.run(max_steps=10)
    assert_within_sensitivity(host(s.q), sens, ref.q, "Riemann 128^2 x10")


def test_riemann_graph_replay_matches_eager():
    """A captured step replayed from the graph gives the eager steps' result."""
    from hit_solver_amd.riemann import RiemannConfig3, RiemannSolver

    cfg = RiemannConfig3(nx=96, ny=80, shear_threshold=None, progress_every=10_000)
    a, b = RiemannSolver(cfg), RiemannSolver(cfg)
    a.run(max_steps=6)
    b.run(max_steps=1)
    b.capture()  # capture runs one step while recording
    b.steps += 1
    for _ in range(4):
        b.replay(1)
        torch.cuda.synchronize()
    np.testing.assert_array_equal(host(b.q), host(a.q))


def test_fp32_step_error_vs_fp64_device():
    """FP32 stencils/solves with FP64 reductions against the device's FP64
    path: per-field error of three HIT steps (the comparisons with the FP64
    oracle are in test_gpu_configs.py)."""
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver

    common = dict(nx=64, ny=64, tfinal=1.0, max_steps=3, viscosity=7.5e-4, target_mach=0.25,
                  initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0)
    s64 = HITSolver(HIT2DConfig(**common))
    s32 = HITSolver(HIT2DConfig(dtype="fp32", **common))
    s64.run()
    s32.run()
    assert_fields(host(s32.q), host(s64.q), 1e-5, "FP32 vs FP64 (3 steps)")
