"""HIT diagnostics and output files on the HIP path vs the NumPy oracle.

Tolerances (the device computes in double whatever the state's dtype):
* rho, u, v, p planes: bit-exact (same arithmetic, same order);
* divergence and vorticity planes: max |diff| <= 1e-11 * the larger of
  max |divergence| and max |vorticity| (the
  reference differentiates by FFT, the device by the equivalent circular
  convolution with the periodic differentiation kernel);
* scalar diagnostics: relative 1e-11 (reduction order; absolute on the
  natural scale for the cancellations mass_error, A_K, C_uv, divergence_rms),
  the WENO fraction exactly (same mask);
* a full run_simulation against the oracle loop: relative 1e-8 (the two
  trajectories differ by rounding after a few steps).
"""

import json

import numpy as np
import pytest
import torch

from oracle import hit as OH
from oracle.equations import Gas

pytestmark = pytest.mark.gpu

DEV = "cuda"


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def close(a, b, rtol=1e-11, atol=1e-15):
    return abs(a - b) <= atol + rtol * abs(b)


def assert_dicts(got, want, rtol=1e-11, atol=1e-15, exact=("weno_fraction",), what=""):
    """Relative rtol, except where the value is a cancellation: mass_error
    (a difference of two masses) to 1e-12 of the mass, the bounded ratios
    A_K and C_uv to 1e-12 absolute, divergence_rms (rounding noise for a
    solenoidal field) to 1e-11 of vorticity_rms."""
    assert list(got) == list(want), f"{what}: keys differ"
    absolute = {"mass_error": 1e-12 * abs(want["mass"]), "A_K": 1e-12, "C_uv": 1e-12,
                "divergence_rms": 1e-11 * want["vorticity_rms"]}
    for k in want:
        if k in absolute:
            assert abs(got[k] - want[k]) <= absolute[k], f"{what} {k}: {got[k]!r} vs {want[k]!r}"
        elif k == "weno_fraction" and k not in exact:
            assert abs(got[k] - want[k]) <= 0.05, f"{what} {k}: {got[k]!r} vs {want[k]!r}"
        elif k in exact:
            assert got[k] == want[k], f"{what} {k}: {got[k]!r} != {want[k]!r}"
        else:
            assert close(got[k], want[k], rtol, atol), f"{what} {k}: {got[k]!r} vs {want[k]!r}"


def smooth_state(nx, ny, seed, noise=0.0):
    """Random smooth state (modes up to |k| = 6) plus optional white noise."""
    rng = np.random.default_rng(seed)
    L = 2.0 * np.pi
    X, Y = np.meshgrid(L / nx * (np.arange(nx) + 0.5), L / ny * (np.arange(ny) + 0.5), indexing="xy")
    f = []
    for _ in range(4):
        a = np.zeros_like(X)
        for _m in range(8):
            kx, ky = rng.integers(-6, 7, 2)
            a += rng.normal() * np.cos(kx * X + ky * Y + rng.uniform(0, 2 * np.pi))
        f.append(a / 8 + noise * rng.normal(size=X.shape))
    rho = 1.0 + 0.2 * f[0]
    u, v = 0.3 * f[1], 0.3 * f[2]
    p = 0.7 + 0.1 * f[3]
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])


def domain_of(nx, ny):
    from hit_solver_amd.domain import Domain2D

    return Domain2D(0.0, 2.0 * np.pi, 0.0, 2.0 * np.pi, nx, ny)


def hit_init(n, seed=1234):
    p = OH.HITParams(nx=n, ny=n, target_mach=0.25, initial_kmin=3, initial_kmax=5)
    return OH.initial_state(p, np.random.default_rng(seed))


CASES = [
    ("hit64_fp64", lambda: hit_init(64), torch.float64),
    ("hit64_fp32", lambda: hit_init(64), torch.float32),
    ("smooth45x38", lambda: smooth_state(45, 38, 3), torch.float64),
    ("noisy40x33", lambda: smooth_state(40, 33, 4, noise=0.05), torch.float64),
]


@pytest.mark.parametrize("name,make,dtype", CASES, ids=[c[0] for c in CASES])
def test_fields_and_diagnostics_vs_oracle(name, make, dtype, tmp_path):
    from hit_solver_amd import hit2d
    from hit_solver_amd.equations import EulerEquation2D

    print(f"[diag] {name}", flush=True)
    q = make()
    ny, nx = q.shape[1:]
    qd = dev(q, dtype)
    qh = qd.double().cpu().numpy()  # what the device sees
    d = domain_of(nx, ny)
    eq = EulerEquation2D(gamma=1.4)
    gas = Gas(gamma=1.4)
    L = 2.0 * np.pi
    # snapshot fields through the file writer
    path = hit2d.save_snapshot(tmp_path, 7, 0.25, qd, eq, d)
    assert path.name == "hit2d_step0000007.npz"
    want = OH.snapshot_fields(qh, gas, nx, ny, d.dx, d.dy)
    with np.load(path, allow_pickle=False) as z:
        assert set(z.files) == {"step", "time", "x", "y", "rho", "u", "v", "pressure", "vorticity", "divergence"}
        assert int(z["step"]) == 7 and float(z["time"]) == 0.25
        np.testing.assert_array_equal(z["x"], d.x)
        for k in ("rho", "u", "v", "pressure"):
            np.testing.assert_array_equal(z[k], want[k], err_msg=k)
        # velocity-gradient scale: the divergence of a solenoidal field is
        # itself rounding noise
        scale = max(np.abs(want["vorticity"]).max(), np.abs(want["divergence"]).max())
        for k in ("vorticity", "divergence"):
            err = np.abs(z[k] - want[k]).max() / scale
            assert err <= 1e-11, f"{k}: {err:.3e}"
    # scalar diagnostics
    info = {"injected_power": 1.1e-3, "Fxx": 0.5, "Fyy": 0.4, "Fxy": 0.01, "A_F": 0.1, "alpha": 0.9,
            "target_power": 1e-3, "mach_control_mach": 0.24, "mach_control_target": 0.25,
            "mach_control_power_desired": 1.2e-3}
    got = hit2d.compute_diagnostics(qd, eq, d, initial_mass=39.0, forcing_info=info, drag_power=-1e-4,
                                    cooling_power=2e-5, large_scale_fraction=0.3, weno_fraction=0.01)
    ref = OH.compute_diagnostics(qh, gas, nx, ny, d.dx, d.dy, L, L, initial_mass=39.0, forcing_info=info,
                                 drag_power=-1e-4, cooling_power=2e-5, large_scale_fraction=0.3,
                                 weno_fraction=0.01)
    assert_dicts(got, ref, what=name)
    for kmax in (1.0, 1.5, 2.0, 3.0, 4.2):
        a = hit2d.large_scale_kinetic_fraction(qd, eq, d, kmax)
        b = OH.large_scale_kinetic_fraction(qh, gas, nx, ny, d.dx, d.dy, kmax)
        assert close(a, b, 1e-10, 1e-15), f"fraction kmax={kmax}: {a} vs {b}"
    # drag power through the plan-level sums
    from hit_solver_amd import _lib

    plan = hit2d._diag_plan(qd, eq, d)
    for kmax, coef in ((2.0, 0.1), (3.0, 0.25)):
        s = plan.diagnostics(qd, lowpass_modes=hit2d.lowpass_modes(d, kmax), drag_coef=coef)
        a = float(s[_lib.DIAG_DRAG_E] / (nx * ny))
        b = OH.source_power(OH.drag_source(qh, gas, nx, ny, d.dx, d.dy, coef, kmax))
        assert close(a, b, 1e-10, 1e-17), f"drag power kmax={kmax}: {a} vs {b}"


def test_divergence_and_vorticity_api():
    from hit_solver_amd import hit2d

    print("[diag] divergence_and_vorticity 48x40", flush=True)
    q = smooth_state(48, 40, 9, noise=0.02)
    u, v = q[1] / q[0], q[2] / q[0]
    d = domain_of(48, 40)
    dv, w = hit2d.divergence_and_vorticity(dev(u), dev(v), d)
    a, b = OH.divergence_and_vorticity(u, v, 48, 40, d.dx, d.dy)
    scale = max(np.abs(a).max(), np.abs(b).max())
    assert np.abs(dv.cpu().numpy() - a).max() <= 1e-11 * scale
    assert np.abs(w.cpu().numpy() - b).max() <= 1e-11 * scale


def test_lowpass_band_limits_are_checked():
    from hit_solver_amd import hit2d
    from hit_solver_amd.equations import EulerEquation2D

    d = domain_of(32, 32)
    qd = dev(smooth_state(32, 32, 1))
    with pytest.raises(ValueError, match="n_lowpass must be in 0..64"):  # 80 modes
        hit2d.large_scale_kinetic_fraction(qd, EulerEquation2D(), d, 5.0)
    plan = hit2d._diag_plan(qd, EulerEquation2D(), d)
    with pytest.raises(ValueError, match="low-pass mode index beyond 4"):
        plan.diagnostics(qd, lowpass_modes=[(5, 0), (-5, 0)])
    with pytest.raises(ValueError, match="excludes k = 0"):
        plan.diagnostics(qd, lowpass_modes=[(0, 0)])


def production(n, **kw):
    base = dict(nx=n, ny=n, cfl=0.05, gamma=1.4, target_mach=0.25, initial_kmin=3, initial_kmax=5,
                forcing_kmin=3.0, forcing_kmax=5.0, target_energy_injection=1.0e-3, forcing_correlation_time=0.5,
                forcing_alpha_memory=0.2, min_forcing_power=1.0e-6, max_forcing_rescale=20.0, mach_control=True,
                mach_control_target=0.25, mach_control_memory=0.995, mach_control_exponent=2.0,
                viscosity=7.5e-4, hyperviscosity_mn=0.002, large_scale_drag=0.10, large_scale_drag_kmax=2.0,
                cooling_time=5.0, forcing_seed=1234, tfinal=1e9)
    base.update(kw)
    return base


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
def test_solver_diagnostics_on_oracle_states(dtype):
    """HITSolver.diagnostics() on the oracle loop's own states (initial, and
    after 6 steps with forcing, drag and cooling) against the oracle loop's
    diagnostics of the same state."""
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver

    print(f"[diag] solver diagnostics 48x48 {dtype}", flush=True)
    run = OH.HITRun(OH.HITParams(**production(48)))
    s = HITSolver(HIT2DConfig(**production(48), dtype=dtype))
    tdt = torch.float64 if dtype == "fp64" else torch.float32
    q0 = run.q if dtype == "fp64" else run.q.astype(np.float32).astype(np.float64)
    s.q = dev(q0, tdt)
    run.q = q0
    # an FP32 plan detects on the FP32 state: its mask may differ near the
    # thresholds, so the WENO fraction is compared to a few dilated nodes
    exact = ("weno_fraction",) if dtype == "fp64" else ()
    first = s.diagnostics(initial=True)
    assert_dicts(first, run.diagnostics(initial=True), exact=exact, what="initial")
    for _ in range(6):
        run.step()
    q6 = run.q if dtype == "fp64" else run.q.astype(np.float32).astype(np.float64)
    run.q = q6
    s.q = dev(q6, tdt)
    s.info = run.info
    got, want = s.diagnostics(initial_mass=first["mass"]), run.diagnostics(initial_mass=first["mass"])
    if dtype == "fp32":
        assert abs(got["weno_fraction"] - want["weno_fraction"]) <= 0.05
        got["weno_fraction"] = want["weno_fraction"]
    assert_dicts(got, want, exact=exact, what="step 6")


REFERENCE_HISTORY_EXTRAS = {
    "forcing_kmin", "forcing_kmax", "forcing_correlation_time", "large_scale_drag", "large_scale_drag_kmax",
    "cooling_time", "cooling_target_pressure", "target_energy_injection", "mach_control", "mach_control_target",
    "mach_control_memory", "mach_control_exponent", "mach_control_min_power", "mach_control_max_power",
}


def test_run_simulation_files_match_oracle_loop(tmp_path):
    """run_simulation writes the reference's files; its history follows the
    oracle loop's diagnostics at the same steps."""
    from hit_solver_amd.hit2d import HIT2DConfig, run_simulation

    print("[diag] run_simulation 32x32, 4 steps", flush=True)
    kw = production(32, max_steps=4)
    cfg = HIT2DConfig(**kw, diagnostics_every=2, snapshot_every=2, output_dir=tmp_path)
    q, t, steps = run_simulation(cfg)
    assert steps == 4
    conf = json.loads((tmp_path / "config.json").read_text())
    assert conf["problem"] == "hit2d" and conf["forcing_type"] == "solenoidal_shell_ou"
    assert conf["spatial_discretization"] == "hybrid_compact_weno7" and conf["nx"] == 32
    for s in (0, 2, 4):
        assert (tmp_path / f"hit2d_step{s:07d}.npz").exists()
    # oracle loop with the same sampling
    p = OH.HITParams(**kw)
    run = OH.HITRun(p)
    want = [dict(step=0.0, time=0.0, dt=0.0, **run.diagnostics(initial=True))]
    m0 = want[0]["mass"]
    while not run.done():
        run.step()
        if run.steps % 2 == 0:
            want.append(dict(step=float(run.steps), time=run.time, dt=run.last_dt,
                             **run.diagnostics(initial_mass=m0)))
    with np.load(tmp_path / "diagnostic_history.npz", allow_pickle=False) as z:
        keys = set(z.files)
        assert keys == set(want[0]) | REFERENCE_HISTORY_EXTRAS
        # the settings written after the records replace the record column of
        # the same name (mach_control_target), as in the reference's writer
        assert float(z["mach_control_target"]) == 0.25
        for k in want[0]:
            if k in REFERENCE_HISTORY_EXTRAS:
                continue
            got = z[k]
            ref = np.array([w[k] for w in want])
            assert got.shape == ref.shape, k
            if k == "weno_fraction":
                assert np.abs(got - ref).max() <= 2.0 / (32 * 32), k
            elif k in ("mass_error", "C_uv", "divergence_rms"):
                assert np.allclose(got, ref, rtol=1e-6, atol=1e-12), f"{k}: {got} vs {ref}"
            else:
                assert np.allclose(got, ref, rtol=1e-8, atol=1e-15), f"{k}: {got} vs {ref}"
        assert float(z["large_scale_drag"]) == 0.10 and bool(z["mach_control"])


# ---------------------------------------------------------------- Riemann outputs
RIEMANN_NPZ_KEYS = {"x", "y", "q", "rho", "u", "v", "pressure", "omega_z", "time", "steps", "configuration_number",
                    "config_json", "diagnostics_json"}


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
def test_riemann_diagnostics_and_files_vs_oracle(dtype, tmp_path):
    """RiemannSolver after 12 steps of configuration 3 at 40x32: the
    diagnostics of its state equal the oracle's _diagnostics of the same
    state (bit for bit: the same arithmetic in double); save_run_outputs
    writes the reference's files and planes."""
    from dataclasses import asdict

    from hit_solver_amd.riemann import RiemannConfig3, RiemannSolver, load_riemann_npz, save_run_outputs
    from oracle import riemann as OR

    print(f"[diag] riemann 40x32 {dtype}", flush=True)
    cfg = RiemannConfig3(nx=40, ny=32, tfinal=0.05, dtype=dtype, progress_every=5)
    s = RiemannSolver(cfg)
    s.run(max_steps=12)
    d = s.diagnostics()
    qh = s.q.double().cpu().numpy()
    p = OR.config3(nx=40, ny=32, tfinal=0.05)
    want = OR.diagnostics(qh, p, s.time, s.steps, s.hv_applications)
    got = asdict(d)
    assert got.pop("backend") == "hip" and want.pop("backend") is None
    if dtype == "fp32":  # the FP32 plan's mask is its own
        assert abs(got.pop("final_sensor_fraction") - want.pop("final_sensor_fraction")) <= 0.1
    assert got == want
    paths = save_run_outputs(s.q, cfg, d, tmp_path, run_id="run1", plot=False)
    run = tmp_path / "run1"
    assert paths["npz"] == run / "run1_final.npz"
    assert json.loads((run / "diagnostics.json").read_text())["steps"] == 12
    conf = json.loads((run / "config.json").read_text())
    assert conf["configuration_number"] == 3 and conf["nx"] == 40
    data, cfg2 = load_riemann_npz(paths["npz"])
    with data:
        assert set(data.files) == RIEMANN_NPZ_KEYS
        f = OR.saved_fields(qh, p)
        for k, v in f.items():
            np.testing.assert_array_equal(data[k], v, err_msg=k)
        np.testing.assert_array_equal(data["q"], qh)
        assert int(data["configuration_number"]) == 3 and int(data["steps"]) == 12
        assert json.loads(str(data["diagnostics_json"]))["density_min"] == d.density_min
    assert cfg2.nx == 40 and cfg2.tfinal == 0.05


def test_riemann_validation_errors():
    """validate_raw_physical_state on states made invalid on the host: a NaN,
    a non-positive density, a non-positive pressure (the reference's messages
    and order)."""
    from hit_solver_amd.equations import EulerEquation2D
    from hit_solver_amd.riemann import RiemannConfig3, validate_raw_physical_state

    print("[diag] riemann validation", flush=True)
    eq = EulerEquation2D()
    q = RiemannConfig3(nx=32, ny=24).initial_state()
    validate_raw_physical_state(q, eq)
    bad = q.clone()
    bad[2, 3, 4] = float("nan")
    with pytest.raises(FloatingPointError, match="contains NaN or Inf values"):
        validate_raw_physical_state(bad, eq, "probe")
    bad = q.clone()
    bad[0, 5, 6] = -1.0
    with pytest.raises(FloatingPointError, match="non-positive density"):
        validate_raw_physical_state(bad, eq)
    bad = q.clone()
    bad[3, 7, 8] = 0.0
    with pytest.raises(FloatingPointError, match="non-positive pressure"):
        validate_raw_physical_state(bad, eq)
    bad = q.clone()
    bad[0, 1, 1] = 0.0  # rho = 0: u = q1/0 -> non-finite pressure
    with pytest.raises(FloatingPointError, match="non-finite density or pressure"):
        validate_raw_physical_state(bad, eq)
