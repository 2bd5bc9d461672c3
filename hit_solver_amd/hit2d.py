"""Forced 2-D HIT on the hip backend: API of reference OOP/hit2d.py.

``HIT2DConfig`` keeps the reference's fields and defaults (OOP/hit2d.py:36-111)
and adds ``dtype`` ("fp64" or "fp32") and ``max_steps``.  ``run_simulation``
runs the reference step loop (OOP/hit2d.py:764-918): dt with the viscous
clip, Mach control, one forcing update, one fused SSP-RK3 step, and the
compact-node hyperviscosity pass every ``hyperviscosity_interval`` steps,
with the reference's diagnostics and files (config.json, hit2d_stepNNNNNNN.npz
snapshots, diagnostic_history.npz; OOP/hit2d.py:429-761) from one device
reduction call per sample.  The initial condition is built on the device
from the phases the reference's generator draws on the host.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from .domain import Domain2D
from .equations import CompressibleNavierStokes2D, EulerEquation2D
from .forcing import IsotropicShellOUForcing2D
from . import _lib
from .runtime import PlanSpec, get_plan
from .spatial_operator import PeriodicEulerShockSensor2D, PeriodicHybridEuler2DOperator, PeriodicHyperviscosity2D
from .time_operator import NativeRHS, SSPRK3


@dataclass(frozen=True)
class HIT2DConfig:
    nx: int = 128
    ny: int = 128
    length: float = 2.0 * np.pi
    gamma: float = 1.4
    cfl: float = 0.05
    tfinal: float = 1.0
    target_mach: float = 0.1
    initial_kmin: int = 1
    initial_kmax: int = 3
    forcing_kmin: float | None = None
    forcing_kmax: float = 3.0
    forcing_rms: float = 1.0
    forcing_correlation_time: float = 1.0
    forcing_alpha_memory: float = 0.2
    target_energy_injection: float | None = 1.0e-3
    min_forcing_power: float = 1.0e-6
    max_forcing_rescale: float | None = 20.0
    mach_control: bool = False
    mach_control_target: float | None = None
    mach_control_memory: float = 0.995
    mach_control_exponent: float = 2.0
    mach_control_min_power: float | None = None
    mach_control_max_power: float | None = None
    forcing_seed: int = 1234
    viscosity: float = 1.0e-3
    prandtl: float = 0.72
    sensor_width: int = 4
    jump_threshold: float = 0.04
    compression_threshold: float = 2.5
    shear_threshold: float | None = None
    hyperviscosity_mn: float = 0.002
    hyperviscosity_interval: int = 5
    large_scale_drag: float = 0.0
    large_scale_drag_kmax: float = 2.0
    cooling_time: float | None = None
    cooling_target_pressure: float | None = None
    diagnostics_every: int = 25
    snapshot_every: int = 100
    forcing_anisotropy_warning: float = 0.2
    flow_anisotropy_warning: float = 0.4
    weno_fraction_warning: float = 0.5
    output_dir: Path = field(default_factory=lambda: Path("hit2d_snapshots"))
    backend: str = "hip"
    dtype: str = "fp64"
    max_steps: int | None = None

    @property
    def torch_dtype(self):
        return torch.float32 if self.dtype == "fp32" else torch.float64

    @classmethod
    def isotropic_128(cls, output_dir: Path = Path("hit2d_snapshots"), backend: str = "hip") -> "HIT2DConfig":
        """Recommended 128x128 annular-shell OU-forced configuration (OOP/hit2d.py:82-111)."""
        return cls(nx=128, ny=128, tfinal=15.0, target_mach=0.25, initial_kmin=3, initial_kmax=5,
                   forcing_kmin=3.0, forcing_kmax=5.0, forcing_correlation_time=0.5, forcing_alpha_memory=0.2,
                   target_energy_injection=1.0e-3, viscosity=7.5e-4, hyperviscosity_mn=0.002,
                   large_scale_drag=0.10, large_scale_drag_kmax=2.0, cooling_time=5.0, snapshot_every=75,
                   diagnostics_every=25, output_dir=output_dir, backend=backend)

    @classmethod
    def production_512(cls, **kw):
        """Production settings of the reference README (README.md:101-129)."""
        base = dict(nx=512, ny=512, cfl=0.05, gamma=1.4, target_mach=0.25, initial_kmin=3, initial_kmax=5,
                    forcing_kmin=3.0, forcing_kmax=5.0, target_energy_injection=1.0e-3,
                    forcing_correlation_time=0.5, forcing_alpha_memory=0.2, min_forcing_power=1.0e-6,
                    max_forcing_rescale=20.0, mach_control=True, mach_control_target=0.25,
                    mach_control_memory=0.995, mach_control_exponent=2.0, viscosity=7.5e-4,
                    hyperviscosity_mn=0.002, large_scale_drag=0.10, large_scale_drag_kmax=2.0,
                    cooling_time=5.0, forcing_seed=1234, tfinal=110.0)
        base.update(kw)
        return cls(**base)


def create_grid(config: HIT2DConfig):
    domain = Domain2D(0.0, config.length, 0.0, config.length, config.nx, config.ny)
    X, Y = domain.mesh()
    return domain, X, Y


def streamfunction_modes(config: HIT2DConfig, rng: np.random.Generator) -> np.ndarray:
    """Rows (kx, ky, amplitude, phase) of the initial streamfunction in the
    reference's loop order, phases drawn from rng as the reference draws them
    (OOP/hit2d.py:180-191): one of each +/- pair, kmin <= |k| <= kmax,
    amplitude 1/|k|^2."""
    rows = []
    K = config.initial_kmax
    for kx in range(-K, K + 1):
        for ky in range(-K, K + 1):
            if ky < 0 or (ky == 0 and kx <= 0):
                continue
            mag = np.sqrt(kx**2 + ky**2)
            if mag < config.initial_kmin or mag > config.initial_kmax:
                continue
            rows.append((kx, ky, 1.0 / max(kx**2 + ky**2, 1), rng.uniform(0.0, 2.0 * np.pi)))
    return np.array(rows, dtype=np.float64).reshape(-1, 4)


def initialize_hit_2d(config: HIT2DConfig, domain: Domain2D, equation, rng: np.random.Generator,
                      dtype=torch.float64, out=None):
    """OOP/hit2d.py:158-204 on the device (hitx_hit_initial_state): the
    streamfunction of streamfunction_modes at the cell centres, u = dpsi/dy,
    v = -dpsi/dx (spectral), mean removal, rho = 1, p = 1/gamma, velocity
    scaled to the target turbulent Mach number.  Returns the conservative
    state [4, ny, nx] of dtype on the device (formed in double)."""
    modes = np.ascontiguousarray(streamfunction_modes(config, rng))
    if len(modes) == 0:
        raise ValueError("initial streamfunction produced zero velocity")
    plan = get_plan(PlanSpec(nx=domain.nx, ny=domain.ny, dx=domain.dx, dy=domain.dy, periodic=True, dtype=dtype,
                             gamma=equation.gamma, rho_floor=equation.rho_floor, p_floor=equation.pressure_floor))
    q = torch.empty((4, domain.ny, domain.nx), dtype=dtype, device="cuda") if out is None else out
    rms = plan.hit_initial_state(modes, config.target_mach, q)
    if rms <= np.finfo(float).eps:
        raise ValueError("initial streamfunction produced zero velocity")
    return q


def drag_modes(domain: Domain2D, kmax: float):
    """Index pairs of the 0 < |k| <= kmax band (OOP/hit2d.py:352-367 mask)."""
    kxv = 2.0 * np.pi * np.fft.fftfreq(domain.nx, d=domain.dx)
    kyv = 2.0 * np.pi * np.fft.fftfreq(domain.ny, d=domain.dy)
    kx, ky = np.meshgrid(kxv, kyv, indexing="xy")
    k2 = kx**2 + ky**2
    mag = np.sqrt(k2)
    mask = (mag >= 0.0) & (k2 > 0.0) & (mag <= float(kmax))
    iy, ix = np.nonzero(mask)
    fx_idx = np.fft.fftfreq(domain.nx, d=1.0 / domain.nx).round().astype(int)
    fy_idx = np.fft.fftfreq(domain.ny, d=1.0 / domain.ny).round().astype(int)
    return np.stack([fx_idx[ix], fy_idx[iy]], axis=1)


def _make_equation(config: HIT2DConfig):
    """OOP/hit2d.py:599-615."""
    if config.viscosity > 0.0:
        return CompressibleNavierStokes2D(gamma=config.gamma, viscosity=config.viscosity, prandtl=config.prandtl)
    return EulerEquation2D(gamma=config.gamma)


def update_mach_controlled_power(config: HIT2DConfig, current_power, current_mach: float):
    """OOP/hit2d.py:293-349 with the turbulent Mach number already reduced on
    the device (sqrt(mean(u^2+v^2)) / mean(c), OOP/hit2d.py:279-290)."""
    target_power = current_power
    target_mach = config.target_mach if config.mach_control_target is None else config.mach_control_target
    desired = np.nan if target_power is None else float(target_power)
    info = {"mach_control_mach": current_mach, "mach_control_target": float(target_mach),
            "mach_control_power_desired": desired}
    if not config.mach_control or target_power is None:
        return target_power, info
    if target_mach <= 0.0:
        raise ValueError("Mach control requires a positive target Mach number")
    if not 0.0 <= config.mach_control_memory < 1.0:
        raise ValueError("mach_control_memory must be in [0, 1)")
    if config.mach_control_exponent <= 0.0:
        raise ValueError("mach_control_exponent must be positive")
    ratio = (target_mach / max(current_mach, 1.0e-8)) ** config.mach_control_exponent
    desired = float(target_power * ratio)
    if config.mach_control_min_power is not None:
        desired = max(desired, config.mach_control_min_power)
    if config.mach_control_max_power is not None:
        desired = min(desired, config.mach_control_max_power)
    nxt = config.mach_control_memory * target_power + (1.0 - config.mach_control_memory) * desired
    info["mach_control_power_desired"] = desired
    return nxt, info


# ---------------------------------------------------------------- diagnostics
def lowpass_modes(domain: Domain2D, kmax: float) -> np.ndarray:
    """(kx, ky) index pairs of the band 0 < |k| <= kmax (OOP/hit2d.py:352-367)."""
    if kmax <= 0.0:
        return np.zeros((0, 2), int)
    return drag_modes(domain, kmax)


def _diag_plan(q, equation, domain: Domain2D):
    if q.shape != (4, domain.ny, domain.nx):
        raise ValueError(f"state shape {tuple(q.shape)} does not match the grid ({domain.ny}, {domain.nx})")
    return get_plan(PlanSpec(nx=domain.nx, ny=domain.ny, dx=domain.dx, dy=domain.dy, periodic=True, dtype=q.dtype,
                             gamma=equation.gamma, rho_floor=equation.rho_floor, p_floor=equation.pressure_floor))


def _kinetic_fraction(s, n: int) -> float:
    """OOP/hit2d.py:443-453 from the device sums."""
    total = float(s[_lib.DIAG_RKE] / n)
    if total <= np.finfo(float).eps:
        return 0.0
    return float(s[_lib.DIAG_LOW] / n) / total


def _cooling_power(s, n: int, equation, cooling_time, target_pressure) -> float:
    """source_power(cooling_source(...)) (OOP/hit2d.py:402-426, 452-456): the
    source is uniform, its mean is the value itself."""
    if cooling_time is None or cooling_time <= 0.0:
        return 0.0
    reference_pressure = 1.0 / equation.gamma if target_pressure is None else target_pressure
    mean_pressure = float(s[_lib.DIAG_P] / n)
    error = (mean_pressure - reference_pressure) / (equation.gamma - 1.0)
    return float(-error / cooling_time)


def _diagnostics_from_sums(s, domain: Domain2D, initial_mass=None, forcing_info=None, drag_power=0.0,
                           cooling_power=0.0, large_scale_fraction=0.0, weno_fraction=0.0) -> dict:
    """The dictionary of OOP/hit2d.py:484-548 from the device sums."""
    n = domain.nx * domain.ny
    mean_rho = float(s[_lib.DIAG_RHO] / n)
    mass = float(mean_rho * (domain.x_max - domain.x_min) * (domain.y_max - domain.y_min))
    uu = float(s[_lib.DIAG_RUU] / n)
    vv = float(s[_lib.DIAG_RVV] / n)
    uv = float(s[_lib.DIAG_RUV] / n)
    component_sum = uu + vv
    covariance_scale = np.sqrt(max(uu * vv, 0.0))
    rms = float(np.sqrt(s[_lib.DIAG_U2V2] / n))
    diagnostics = {
        "kinetic_energy": 0.5 * component_sum,
        "Kx": 0.5 * uu,
        "Ky": 0.5 * vv,
        "A_K": abs(uu - vv) / component_sum if component_sum > 0.0 else 0.0,
        "C_uv": uv / covariance_scale if covariance_scale > 0.0 else 0.0,
        "rms_velocity": rms,
        "turbulent_mach": float(rms / (s[_lib.DIAG_C] / n)),
        "mean_density": mean_rho,
        "mean_pressure": float(s[_lib.DIAG_P] / n),
        "divergence_rms": float(np.sqrt(s[_lib.DIAG_DIV2] / n)),
        "vorticity_rms": float(np.sqrt(s[_lib.DIAG_VORT2] / n)),
        "mass": mass,
        "mass_error": 0.0 if initial_mass is None else mass - initial_mass,
        "large_scale_kinetic_fraction": large_scale_fraction,
        "drag_power": drag_power,
        "cooling_power": cooling_power,
        "weno_fraction": weno_fraction,
    }
    diagnostics.update({
        "P_in": 0.0, "Fxx": 0.0, "Fyy": 0.0, "Fxy": 0.0, "A_F": 0.0, "forcing_alpha": 0.0,
        "forcing_target_power": 0.0, "mach_control_mach": diagnostics["turbulent_mach"],
        "mach_control_target": 0.0, "mach_control_power_desired": 0.0,
    })
    if forcing_info is not None:
        diagnostics.update({
            "P_in": forcing_info["injected_power"],
            "Fxx": forcing_info["Fxx"],
            "Fyy": forcing_info["Fyy"],
            "Fxy": forcing_info["Fxy"],
            "A_F": forcing_info["A_F"],
            "forcing_alpha": forcing_info["alpha"],
            "forcing_target_power": forcing_info.get("target_power", 0.0),
            "mach_control_mach": forcing_info.get("mach_control_mach", diagnostics["turbulent_mach"]),
            "mach_control_target": forcing_info.get("mach_control_target", 0.0),
            "mach_control_power_desired": forcing_info.get("mach_control_power_desired", 0.0),
        })
    return diagnostics


def compute_diagnostics(q, equation, domain: Domain2D, initial_mass=None, forcing_info=None, drag_power=0.0,
                        cooling_power=0.0, large_scale_fraction=0.0, weno_fraction=0.0) -> dict:
    """OOP/hit2d.py:471-548 for a device state q [4, ny, nx]; sums in double on
    the device (hitx_diagnostics)."""
    s = _diag_plan(q, equation, domain).diagnostics(q)
    return _diagnostics_from_sums(s, domain, initial_mass, forcing_info, drag_power, cooling_power,
                                  large_scale_fraction, weno_fraction)


def large_scale_kinetic_fraction(q, equation, domain: Domain2D, kmax: float) -> float:
    """OOP/hit2d.py:429-453: fraction of the fluctuation kinetic energy in the
    band 0 < |k| <= kmax (|kx|, |ky| <= 4 on this backend)."""
    if kmax <= 0.0:
        return 0.0
    s = _diag_plan(q, equation, domain).diagnostics(q, lowpass_modes=lowpass_modes(domain, kmax))
    return _kinetic_fraction(s, domain.nx * domain.ny)


def divergence_and_vorticity(u, v, domain: Domain2D):
    """OOP/hit2d.py:459-468: spectral divergence and vorticity of device
    velocity planes [ny, nx] (float64 results)."""
    if u.shape != (domain.ny, domain.nx) or v.shape != u.shape:
        raise ValueError("u and v must be [ny, nx] planes of the grid")
    u64, v64 = u.to(torch.float64), v.to(torch.float64)
    one = torch.ones_like(u64)
    q = torch.stack([one, u64, v64, 0.5 * (u64 * u64 + v64 * v64) + 1.0]).contiguous()
    _s, planes = _diag_plan(q, EulerEquation2D(), domain).diagnostics(q, fields=True)
    return planes[4], planes[5]


# ---------------------------------------------------------------- files
def _json_dict(instance) -> dict:
    """Dataclass fields as JSON values (OOP/run_utils.py:27-37)."""
    from dataclasses import asdict

    out = {}
    for key, value in asdict(instance).items():
        if isinstance(value, Path):
            out[key] = str(value)
        elif isinstance(value, (np.integer, np.floating)):
            out[key] = value.item()
        else:
            out[key] = value
    return out


def save_run_config(output_dir: Path, config: HIT2DConfig) -> Path:
    """config.json with the resolved configuration (OOP/hit2d.py:551-568)."""
    import json

    payload = _json_dict(config)
    payload.update({
        "problem": "hit2d",
        "forcing_type": "disabled" if config.forcing_rms <= 0.0 else "solenoidal_shell_ou",
        "spatial_discretization": "hybrid_compact_weno7",
        "hyperviscosity_policy": "compact_nodes_only",
    })
    path = Path(output_dir) / "config.json"
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(payload, indent=2, sort_keys=True))
    return path


def save_snapshot(output_dir: Path, step: int, time: float, q, equation, domain: Domain2D) -> Path:
    """hit2d_step{step:07d}.npz with x, y, rho, u, v, pressure, vorticity,
    divergence (OOP/hit2d.py:571-596); fields from the device in float64."""
    _s, planes = _diag_plan(q, equation, domain).diagnostics(q, fields=True)
    f = planes.cpu().numpy()
    output_dir = Path(output_dir)
    output_dir.mkdir(parents=True, exist_ok=True)
    path = output_dir / f"hit2d_step{step:07d}.npz"
    np.savez_compressed(path, step=step, time=time, x=domain.x, y=domain.y, rho=f[0], u=f[1], v=f[2],
                        pressure=f[3], vorticity=f[5], divergence=f[4])
    return path


def save_diagnostic_history(output_dir: Path, history: list, config: HIT2DConfig) -> Path:
    """diagnostic_history.npz: one array per diagnostic plus the forcing and
    control settings (OOP/hit2d.py:640-694)."""
    output_dir = Path(output_dir)
    output_dir.mkdir(parents=True, exist_ok=True)
    path = output_dir / "diagnostic_history.npz"
    data = {key: np.asarray([record[key] for record in history]) for key in tuple(history[0])}
    nan_if_none = lambda v: np.asarray(np.nan if v is None else v)  # noqa: E731
    data.update({
        "forcing_kmin": np.asarray(0.0 if config.forcing_kmin is None else config.forcing_kmin),
        "forcing_kmax": np.asarray(config.forcing_kmax),
        "forcing_correlation_time": np.asarray(config.forcing_correlation_time),
        "large_scale_drag": np.asarray(config.large_scale_drag),
        "large_scale_drag_kmax": np.asarray(config.large_scale_drag_kmax),
        "cooling_time": nan_if_none(config.cooling_time),
        "cooling_target_pressure": nan_if_none(config.cooling_target_pressure),
        "target_energy_injection": nan_if_none(config.target_energy_injection),
        "mach_control": np.asarray(config.mach_control),
        "mach_control_target": nan_if_none(config.mach_control_target),
        "mach_control_memory": np.asarray(config.mach_control_memory),
        "mach_control_exponent": np.asarray(config.mach_control_exponent),
        "mach_control_min_power": nan_if_none(config.mach_control_min_power),
        "mach_control_max_power": nan_if_none(config.mach_control_max_power),
    })
    np.savez_compressed(path, **data)
    return path


def _diagnostic_record(step: int, time: float, dt: float, diagnostics: dict) -> dict:
    return {"step": float(step), "time": time, "dt": dt, **diagnostics}


def _print_diagnostics(step: int, time: float, dt: float, d: dict) -> None:
    """The progress line of OOP/hit2d.py:618-637."""
    print(
        f"step={step:6d}, t={time:.6f}, dt={dt:.3e}, "
        f"KE={d['kinetic_energy']:.6e}, Mt={d['turbulent_mach']:.4f}, A_K={d['A_K']:.3f}, "
        f"Cuv={d['C_uv']:.3f}, P_in={d['P_in']:.3e}, P_tgt={d['forcing_target_power']:.3e}, "
        f"A_F={d['A_F']:.3f}, K_low={d['large_scale_kinetic_fraction']:.3f}, P_drag={d['drag_power']:.3e}, "
        f"P_cool={d['cooling_power']:.3e}, WENO={d['weno_fraction']:.3f}, rho_mean={d['mean_density']:.6f}, "
        f"p_mean={d['mean_pressure']:.6f}, div_rms={d['divergence_rms']:.3e}, "
        f"vort_rms={d['vorticity_rms']:.3e}, mass_err={d['mass_error']:.3e}"
    )


def _warn_diagnostic_state(diagnostics: dict, history: list, config: HIT2DConfig) -> None:
    """Non-finite check and anisotropy/WENO warnings (OOP/hit2d.py:711-761)."""
    if not all(np.isfinite(value) for value in diagnostics.values()):
        raise FloatingPointError("NaN or infinite value detected in HIT2D diagnostics")
    if history:
        fxx = float(np.mean([item["Fxx"] for item in history]))
        fyy = float(np.mean([item["Fyy"] for item in history]))
        fxy = float(np.mean([item["Fxy"] for item in history]))
        force_sum = fxx + fyy
        cumulative_af = abs(fxx - fyy) / force_sum if force_sum > 0.0 else 0.0
        normalized_fxy = abs(fxy) / force_sum if force_sum > 0.0 else 0.0
        if cumulative_af > config.forcing_anisotropy_warning:
            print(f"warning: cumulative forcing anisotropy is high (A_F={cumulative_af:.3f})")
        if normalized_fxy > 0.1:
            print(f"warning: cumulative forcing cross-correlation is high (|Fxy|/(Fxx+Fyy)={normalized_fxy:.3f})")
    if diagnostics["A_K"] > config.flow_anisotropy_warning:
        print(f"warning: flow component anisotropy is high (A_K={diagnostics['A_K']:.3f})")
    if len(history) >= 3:
        recent = [item["A_K"] for item in history[-3:]]
        if recent[0] < recent[1] < recent[2] and recent[-1] > 0.2:
            print("warning: flow component anisotropy has increased over the last three diagnostic samples")
    if diagnostics["weno_fraction"] > config.weno_fraction_warning:
        print(f"warning: WENO is active over a large domain fraction ({diagnostics['weno_fraction']:.3f})")
    if diagnostics["large_scale_kinetic_fraction"] > 0.35 and config.large_scale_drag <= 0.0:
        print(
            "warning: a large fraction of kinetic energy is in low-k modes "
            f"(K_low={diagnostics['large_scale_kinetic_fraction']:.3f}); consider enabling --large-scale-drag"
        )


class HITSolver:
    """The reference HIT step loop as an object; ``step()`` advances one step."""

    def __init__(self, config: HIT2DConfig, q0=None):
        if config.backend not in ("hip", "auto"):
            raise ValueError("this package implements the 'hip' backend")
        if config.hyperviscosity_interval <= 0 and config.hyperviscosity_mn > 0.0:
            raise ValueError("hyperviscosity_interval must be positive when mn > 0")
        self.config = c = config
        self.dtype = c.torch_dtype
        self.domain, _, _ = create_grid(c)
        d = self.domain
        self.equation = _make_equation(c)
        self.sensor = PeriodicEulerShockSensor2D(d, self.equation, width=c.sensor_width,
                                                 compression_threshold=c.compression_threshold,
                                                 jump_threshold=c.jump_threshold, shear_threshold=c.shear_threshold)
        self.operator = PeriodicHybridEuler2DOperator(d, self.equation, self.sensor)
        self.hyperviscosity = PeriodicHyperviscosity2D(d, mn=c.hyperviscosity_mn)
        self.integrator = SSPRK3()
        rng = np.random.default_rng(c.forcing_seed)
        self.forcing = IsotropicShellOUForcing2D(
            domain=d, k_min=0.0 if c.forcing_kmin is None else c.forcing_kmin, k_max=c.forcing_kmax,
            correlation_time=c.forcing_correlation_time, force_rms=c.forcing_rms,
            target_power=c.target_energy_injection, min_power=c.min_forcing_power,
            max_rescale=c.max_forcing_rescale, alpha_memory=c.forcing_alpha_memory, seed=c.forcing_seed + 1,
            dtype=self.dtype)
        if q0 is None:
            q0 = initialize_hit_2d(c, d, self.equation, rng, dtype=self.dtype)
        eq = self.equation
        self.q = eq.enforce_physical_state(q0)
        dm = drag_modes(d, c.large_scale_drag_kmax) if c.large_scale_drag > 0.0 and c.large_scale_drag_kmax > 0.0 \
            else np.zeros((0, 2), int)
        sources = dict(
            viscosity=c.viscosity, prandtl=c.prandtl, forcing=True,
            drag_coef=c.large_scale_drag if len(dm) else 0.0, drag_modes=tuple(int(v) for v in dm.reshape(-1)),
            cooling_time=c.cooling_time if (c.cooling_time is not None and c.cooling_time > 0.0) else None,
            cooling_p_ref=c.cooling_target_pressure)
        self.stage_plan = get_plan(self.sensor.plan_spec(self.dtype, **sources))
        self.step_plan = get_plan(self.sensor.plan_spec(self.dtype))
        self.hv_plan = get_plan(self.sensor.plan_spec(self.dtype, hv_coef=self.hyperviscosity.coefficients()))
        self.time = 0.0
        self.steps = 0
        self.power = c.target_energy_injection
        self.info = None
        self.last_dt = None
        self._spare = torch.empty_like(self.q)
        self._mask = torch.empty(self.q.shape[1:], dtype=torch.uint8, device=self.q.device)

    def done(self):
        c = self.config
        if c.max_steps is not None and self.steps >= c.max_steps:
            return True
        return not self.time < c.tfinal

    def step(self):
        c, d, q = self.config, self.domain, self.q
        max_sr, s_uv, s_c = self.step_plan.step_reduce(q)
        dt = c.cfl / max_sr
        dt = min(dt, c.tfinal - self.time)
        if c.viscosity > 0.0:
            h = min(d.dx, d.dy)
            dt = min(dt, 0.25 * h**2 / c.viscosity, c.tfinal - self.time)
        N = float(d.nx * d.ny)
        mach = float(np.sqrt(s_uv / N) / (s_c / N))
        self.power, minfo = update_mach_controlled_power(c, self.power, mach)
        self.forcing.target_power = self.power
        rho, u, v, _p = self.step_plan.primitives(q)
        fx, fy, info = self.forcing.update(dt, rho, u, v)
        info.update(minfo)
        self.info = info
        rhs = NativeRHS(self.stage_plan, fx, fy)
        q_new = self.integrator.step(q, rhs, dt, clean=rhs.clean, out=self._spare)
        self._spare = q
        self.steps += 1
        self.time += dt
        self.last_dt = dt
        if c.hyperviscosity_mn > 0.0 and self.steps % c.hyperviscosity_interval == 0:
            weno = self.step_plan.sensor(q_new, enforce_first=False, out=self._mask)
            q_hv = self.hv_plan.hyperviscosity(q_new, weno, apply_guard=False, out=self._spare, invert=True)
            self._spare = q_new
            q_new = q_hv
        self.q = q_new
        return q_new

    def diagnostics(self, initial: bool = False, initial_mass=None) -> dict:
        """The diagnostics of the current state as the reference loop forms
        them (OOP/hit2d.py:806-818 initial, :881-905 in the loop): WENO
        fraction of detect(q), drag and cooling source powers, large-scale
        kinetic fraction, and the forcing information of the last step.  One
        sensor launch and one hitx_diagnostics call."""
        c, d = self.config, self.domain
        n = d.nx * d.ny
        mask = self.step_plan.sensor(self.q, enforce_first=False, out=self._mask)
        kmax = c.large_scale_drag_kmax
        drag_on = c.large_scale_drag > 0.0 and kmax > 0.0
        s = self.step_plan.diagnostics(self.q, mask, lowpass_modes(d, kmax),
                                       c.large_scale_drag if drag_on else 0.0)
        fraction = _kinetic_fraction(s, n) if kmax > 0.0 else 0.0
        weno = float(s[_lib.DIAG_MASK] / n)
        if initial:
            return _diagnostics_from_sums(s, d, large_scale_fraction=fraction, weno_fraction=weno)
        drag_power = float(s[_lib.DIAG_DRAG_E] / n) if drag_on else 0.0
        cooling_power = _cooling_power(s, n, self.equation, c.cooling_time, c.cooling_target_pressure)
        return _diagnostics_from_sums(s, d, initial_mass, self.info, drag_power, cooling_power, fraction, weno)

    def run(self):
        while not self.done():
            self.step()
        return self.q, self.time, self.steps


def run_simulation(config: HIT2DConfig):
    """OOP/hit2d.py:764-918 with its files: config.json, the step-0 snapshot
    and diagnostics, a diagnostics sample every ``diagnostics_every`` steps
    and a snapshot every ``snapshot_every`` steps (and both at the end).
    Returns (q, time, steps)."""
    print("Using array backend: hip")
    if config.hyperviscosity_interval <= 0 and config.hyperviscosity_mn > 0.0:
        raise ValueError("hyperviscosity_interval must be positive when mn > 0")
    config_path = save_run_config(config.output_dir, config)
    print(f"saved run configuration: {config_path}")
    solver = HITSolver(config)
    first = solver.diagnostics(initial=True)
    initial_mass = first["mass"]
    history = [_diagnostic_record(0, 0.0, 0.0, first)]
    save_snapshot(config.output_dir, 0, 0.0, solver.q, solver.equation, solver.domain)
    save_diagnostic_history(config.output_dir, history, config)
    _print_diagnostics(0, 0.0, 0.0, first)
    while not solver.done():
        solver.step()
        step, time, dt = solver.steps, solver.time, solver.last_dt
        last = time >= config.tfinal or solver.done()
        if step % config.diagnostics_every == 0 or last:
            diag = solver.diagnostics(initial_mass=initial_mass)
            history.append(_diagnostic_record(step, time, dt, diag))
            _warn_diagnostic_state(diag, history[1:], config)
            save_diagnostic_history(config.output_dir, history, config)
            _print_diagnostics(step, time, dt, diag)
        if step % config.snapshot_every == 0 or last:
            path = save_snapshot(config.output_dir, step, time, solver.q, solver.equation, solver.domain)
            print(f"saved snapshot: {path}")
    return solver.q, solver.time, solver.steps
