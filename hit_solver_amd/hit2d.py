"""Forced 2-D HIT on the hip backend: API of reference OOP/hit2d.py.

``HIT2DConfig`` keeps the reference's fields and defaults (OOP/hit2d.py:36-111)
and adds ``dtype`` ("fp64" or "fp32") and ``max_steps``.  ``run_simulation``
runs the reference step loop (OOP/hit2d.py:764-918): dt with the viscous
clip, Mach control, one forcing update, one fused SSP-RK3 step, and the
compact-node hyperviscosity pass every ``hyperviscosity_interval`` steps.
Diagnostics and file output (OOP/hit2d.py:429-761) are not part of this
path.  The initial condition is built on the host (setup, once).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from .domain import Domain2D
from .equations import CompressibleNavierStokes2D, EulerEquation2D
from .forcing import IsotropicShellOUForcing2D
from .runtime import get_plan
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
    output_dir: Path = field(default_factory=lambda: Path("hit2d_snapshots"))
    backend: str = "hip"
    dtype: str = "fp64"
    max_steps: int | None = None

    @property
    def torch_dtype(self):
        return torch.float32 if self.dtype == "fp32" else torch.float64

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


def initialize_hit_2d(config: HIT2DConfig, domain: Domain2D, equation, rng: np.random.Generator):
    """OOP/hit2d.py:158-204 (host, setup): random-phase streamfunction on the
    half-plane of the shell, amplitude 1/k^2, spectral curl, mean removal,
    rho = 1, p = 1/gamma, velocity scaled to the target turbulent Mach number.
    Returns the conservative state as float64 numpy [4, ny, nx]."""
    X, Y = domain.mesh()
    psi = np.zeros_like(X)
    K = config.initial_kmax
    for kx in range(-K, K + 1):
        for ky in range(-K, K + 1):
            if ky < 0 or (ky == 0 and kx <= 0):
                continue
            mag = np.sqrt(kx**2 + ky**2)
            if mag < config.initial_kmin or mag > config.initial_kmax:
                continue
            amp = 1.0 / max(kx**2 + ky**2, 1)
            phase = rng.uniform(0.0, 2.0 * np.pi)
            psi += amp * np.cos(kx * X + ky * Y + phase)
    kxv = 2.0 * np.pi * np.fft.fftfreq(domain.nx, d=domain.dx)
    kyv = 2.0 * np.pi * np.fft.fftfreq(domain.ny, d=domain.dy)
    KX, KY = np.meshgrid(kxv, kyv, indexing="xy")
    ph = np.fft.fft2(psi)
    u = np.fft.ifft2(1j * KY * ph).real
    v = -np.fft.ifft2(1j * KX * ph).real
    u -= np.mean(u)
    v -= np.mean(v)
    rho = np.ones_like(u)
    p = np.full_like(u, 1.0 / config.gamma)
    c = np.sqrt(equation.gamma * p / rho)
    vrms = np.sqrt(np.mean(u**2 + v**2))
    if float(vrms) <= np.finfo(float).eps:
        raise ValueError("initial streamfunction produced zero velocity")
    scale = config.target_mach * float(np.mean(c)) / vrms
    u *= scale
    v *= scale
    rho = np.maximum(rho, equation.rho_floor)
    p = np.maximum(p, equation.pressure_floor)
    q = np.zeros((4, domain.ny, domain.nx))
    q[0] = rho
    q[1] = rho * u
    q[2] = rho * v
    q[3] = p / (equation.gamma - 1.0) + 0.5 * rho * (u**2 + v**2)
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
            q0 = torch.from_numpy(initialize_hit_2d(c, d, self.equation, rng)).to("cuda", self.dtype)
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

    def run(self):
        while not self.done():
            self.step()
        return self.q, self.time, self.steps


def run_simulation(config: HIT2DConfig):
    """OOP/hit2d.py:764-918 without file output: returns (q, time, steps)."""
    return HITSolver(config).run()
