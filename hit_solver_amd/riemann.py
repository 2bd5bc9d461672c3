"""Lax-Liu 2-D Riemann problems on the hip backend.

API of reference OOP/problems/riemann_config3.py (RiemannConfig3,
NonPeriodicEulerShockSensor2D, NonPeriodicHybridEuler2DOperator,
LocalHyperviscosity2D, apply_outflow_guard, validate_raw_physical_state,
hyperviscosity_enabled_for_config, run_riemann), riemann_config3_08.py and
riemann_config6.py.  The operator, sensor, hyperviscosity and guard run as
HIP kernels, and so does the quadrant initial state.
"""

from __future__ import annotations

import json
import math
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import ClassVar

import numpy as np
import torch

from . import _lib
from .domain import Domain2D
from .equations import EulerEquation2D
from .runtime import PlanSpec, get_plan
from .time_operator import NativeRHS, SSPRK3


@dataclass(frozen=True)
class RiemannConfig3:
    """Lax-Liu / Kurganov-Tadmor configuration 3 (riemann_config3.py:19-58)."""

    configuration_number: ClassVar[int] = 3

    nx: int = 512
    ny: int = 512
    tfinal: float = 0.3
    cfl: float = 0.4
    gamma: float = 1.4
    x_min: float = 0.0
    x_max: float = 1.0
    y_min: float = 0.0
    y_max: float = 1.0
    x_split: float = 0.5
    y_split: float = 0.5
    backend: str = "hip"
    scheme: str = "hybrid"
    sensor_width: int = 4
    jump_threshold: float = 0.025
    compression_threshold: float = 2.5
    shear_threshold: float | None = 0.0
    boundary_guard: int = 4
    mn: float = 0.001
    hyperviscosity_interval: int = 1
    hyperviscosity_density_weight: float = 1.0
    hyperviscosity_momentum_weight: float = 1.0
    hyperviscosity_energy_weight: float = 1.0
    progress_every: int = 500
    guard_cells: int = 2
    save_final_npz: bool = True
    dtype: str = "fp64"

    # (rho, u, v, p): I top-right, II top-left, III bottom-left, IV bottom-right
    quadrants: ClassVar[tuple] = ((1.5, 0.0, 0.0, 1.5), (0.5323, 1.206, 0.0, 0.3),
                                  (0.138, 1.206, 1.206, 0.029), (0.5323, 0.0, 1.206, 0.3))

    @property
    def domain(self) -> Domain2D:
        return Domain2D(self.x_min, self.x_max, self.y_min, self.y_max, self.nx, self.ny)

    @property
    def equation(self) -> EulerEquation2D:
        return EulerEquation2D(gamma=self.gamma)

    @property
    def torch_dtype(self):
        return torch.float32 if self.dtype == "fp32" else torch.float64

    def initial_state(self):
        """Conservative quadrant state on the device (hitx_riemann_initial_state),
        formed in double with the floors of OOP/equations.py:147-163."""
        d, eq = self.domain, self.equation
        plan = get_plan(PlanSpec(nx=self.nx, ny=self.ny, dx=d.dx, dy=d.dy, periodic=False, dtype=self.torch_dtype,
                                 gamma=self.gamma, rho_floor=eq.rho_floor, p_floor=eq.pressure_floor))
        q = torch.empty((4, self.ny, self.nx), dtype=self.torch_dtype, device="cuda")
        return plan.riemann_initial_state(self.quadrants, self.x_min, self.y_min, self.x_split, self.y_split, q)


@dataclass(frozen=True)
class RiemannConfig3Offset08(RiemannConfig3):
    """riemann_config3_08.py:9-19."""

    scenario_name: ClassVar[str] = "riemann_config3_08"
    configuration_label: ClassVar[str] = "Configuration 3 (x0=y0=0.8)"
    x_split: float = 0.8
    y_split: float = 0.8
    hyperviscosity_interval: int = 5


@dataclass(frozen=True)
class RiemannConfig6(RiemannConfig3):
    """riemann_config6.py:40-87: four interacting contact discontinuities."""

    configuration_number: ClassVar[int] = 6
    tfinal: float = 0.25
    quadrants: ClassVar[tuple] = ((1.0, 0.75, -0.50, 1.0), (2.0, 0.75, 0.50, 1.0),
                                  (1.0, -0.75, 0.50, 1.0), (3.0, -0.75, -0.50, 1.0))


@dataclass(frozen=True)
class RiemannDiagnostics:
    backend: str
    scheme: str
    hyperviscosity_enabled: bool
    hyperviscosity_applications: int
    final_sensor_fraction: float
    density_min: float
    density_max: float
    pressure_min: float
    pressure_max: float
    max_abs_vorticity: float
    has_nonfinite: bool
    density_positive: bool
    pressure_positive: bool
    time: float
    steps: int


@dataclass
class NonPeriodicLineCompactDerivative:
    """Compact-scheme constants (riemann_config3.py:225-265); the Thomas-type
    solve with identity end rows runs inside the sweep kernels."""

    n: int
    dx: float
    alpha: float = 3.0 / 8.0


@dataclass(frozen=True)
class NonPeriodicEulerShockSensor2D:
    """riemann_config3.py:278-326."""

    domain: Domain2D
    equation: EulerEquation2D
    width: int = 4
    compression_threshold: float = 2.5
    jump_threshold: float = 0.04
    shear_threshold: float | None = 2.0
    boundary_guard: int = 4

    def plan_spec(self, dtype, **extra) -> PlanSpec:
        eq = self.equation
        return PlanSpec(
            nx=self.domain.nx, ny=self.domain.ny, dx=self.domain.dx, dy=self.domain.dy, periodic=False,
            dtype=dtype, gamma=eq.gamma, rho_floor=eq.rho_floor, p_floor=eq.pressure_floor,
            sensor_width=self.width, boundary_guard=self.boundary_guard,
            compression_threshold=self.compression_threshold, jump_threshold=self.jump_threshold,
            shear_threshold=self.shear_threshold, **extra,
        )

    def detect_u8(self, q):
        return get_plan(self.plan_spec(q.dtype)).sensor(q, enforce_first=False)

    def detect(self, q):
        return self.detect_u8(q).bool()


@dataclass
class NonPeriodicHybridEuler2DOperator:
    """riemann_config3.py:329-378."""

    domain: Domain2D
    equation: EulerEquation2D
    sensor: NonPeriodicEulerShockSensor2D
    scheme: str = "hybrid"
    compact_x: NonPeriodicLineCompactDerivative = None
    compact_y: NonPeriodicLineCompactDerivative = None

    def __post_init__(self):
        if self.scheme not in {"hybrid", "weno"}:
            raise ValueError("scheme must be 'hybrid' or 'weno'")
        self.compact_x = NonPeriodicLineCompactDerivative(self.domain.nx, self.domain.dx)
        self.compact_y = NonPeriodicLineCompactDerivative(self.domain.ny, self.domain.dy)

    def plan(self, dtype, **extra):
        return get_plan(self.sensor.plan_spec(dtype, scheme=self.scheme, **extra))

    def rhs(self, q):
        return self.plan(q.dtype).rhs(q)

    def fixed_time_step(self, q, cfl: float, t_end: float):
        max_sr, _, _ = self.plan(q.dtype).step_reduce(q)
        dt = cfl / max_sr
        n = int(math.ceil(t_end / dt))
        return t_end / n, n

    def native_rhs(self, dtype, guard_cells: int):
        """Stage-fused RHS whose clean is enforce + outflow guard."""
        return NativeRHS(self.plan(dtype, guard_cells=guard_cells))


@dataclass(frozen=True)
class LocalHyperviscosity2D:
    """riemann_config3.py:381-416: edge-padded biharmonic filter on compact nodes."""

    domain: Domain2D
    mn: float = 0.001
    density_weight: float = 1.0
    momentum_weight: float = 1.0
    energy_weight: float = 1.0

    def coefficients(self):
        h = min(self.domain.dx, self.domain.dy)
        w = (self.density_weight, self.momentum_weight, self.momentum_weight, self.energy_weight)
        return tuple((self.mn * wk) * h**4 for wk in w)

    def spec(self, dtype, equation, guard_cells=0):
        d = self.domain
        return PlanSpec(nx=d.nx, ny=d.ny, dx=d.dx, dy=d.dy, periodic=False, dtype=dtype, gamma=equation.gamma,
                        rho_floor=equation.rho_floor, p_floor=equation.pressure_floor,
                        hv_coef=self.coefficients(), guard_cells=guard_cells)

    def apply(self, q, compact_mask, equation, guard_cells: int = 0):
        if tuple(compact_mask.shape) != tuple(q.shape[-2:]):
            raise ValueError(f"compact_mask shape {tuple(compact_mask.shape)} does not match grid {tuple(q.shape[-2:])}")
        act = compact_mask if compact_mask.dtype == torch.uint8 else compact_mask.to(torch.uint8)
        plan = get_plan(self.spec(q.dtype, equation, guard_cells))
        return plan.hyperviscosity(q, act.contiguous(), apply_guard=guard_cells > 0)


def apply_outflow_guard(q, guard_cells: int = 2, equation: EulerEquation2D | None = None):
    """riemann_config3.py:419-429.  Runs fused with enforce on the device, so
    it is exact only on states that enforce leaves unchanged (every state the
    solver produces has been enforced)."""
    if guard_cells <= 0:
        return q
    eq = equation or EulerEquation2D()
    ny, nx = q.shape[1:]
    plan = get_plan(PlanSpec(nx=int(nx), ny=int(ny), dx=1.0, dy=1.0, periodic=False, dtype=q.dtype,
                             gamma=eq.gamma, rho_floor=eq.rho_floor, p_floor=eq.pressure_floor,
                             guard_cells=guard_cells))
    return plan.enforce(q, apply_guard=True)


def raw_density_pressure(q, equation):
    """riemann_config3.py:432-439 (elementwise; the checks below run on the
    device through hitx_riemann_diagnostics)."""
    rho = q[0]
    u = q[1] / rho
    v = q[2] / rho
    ke = 0.5 * rho * (u**2 + v**2)
    return rho, (equation.gamma - 1.0) * (q[3] - ke)


def _check_plan(q, equation, domain: Domain2D | None = None):
    ny, nx = (int(n) for n in q.shape[1:])
    dx, dy = (domain.dx, domain.dy) if domain is not None else (1.0, 1.0)
    return get_plan(PlanSpec(nx=nx, ny=ny, dx=dx, dy=dy, periodic=False, dtype=q.dtype, gamma=equation.gamma,
                             rho_floor=equation.rho_floor, p_floor=equation.pressure_floor))


def validate_raw_physical_state(q, equation, label: str = "state") -> None:
    """riemann_config3.py:442-452 (raises FloatingPointError); one device
    reduction."""
    v = _check_plan(q, equation).riemann_diagnostics(q)
    if v[_lib.RDIAG_NONFINITE_Q] > 0:
        raise FloatingPointError(f"{label} contains NaN or Inf values")
    if v[_lib.RDIAG_NONFINITE_RP] > 0:
        raise FloatingPointError(f"{label} has non-finite density or pressure")
    if v[_lib.RDIAG_RHO_NONPOS] > 0:
        raise FloatingPointError(f"{label} contains non-positive density")
    if v[_lib.RDIAG_P_NONPOS] > 0:
        raise FloatingPointError(f"{label} contains non-positive pressure")


def hyperviscosity_enabled_for_config(config: RiemannConfig3) -> bool:
    """riemann_config3.py:667-678."""
    return config.scheme == "hybrid" and config.mn > 0.0 and config.hyperviscosity_interval > 0


def primitive_fields(q, equation, domain: Domain2D | None = None):
    """riemann_config3.py:455-456: rho, u, v, p (floored primitives) as
    float64 device planes."""
    _v, planes = _check_plan(q, equation, domain).riemann_diagnostics(q, fields=True)
    return planes[0], planes[1], planes[2], planes[3]


def vorticity_z(q, domain: Domain2D, equation):
    """riemann_config3.py:459-463: dv/dx - du/dy of the floored primitives,
    central inside and one-sided at the ends (:268-275); float64 plane."""
    _v, planes = _check_plan(q, equation, domain).riemann_diagnostics(q, fields=True)
    return planes[4]


def _diagnostics(q, config: RiemannConfig3, time: float, steps: int, hyperviscosity_applications: int, sensor):
    """riemann_config3.py:629-650 from one sensor launch and one device
    reduction."""
    mask = sensor.detect_u8(q)
    v = _check_plan(q, config.equation, config.domain).riemann_diagnostics(q, mask)
    n = config.nx * config.ny
    return RiemannDiagnostics(
        backend=config.backend, scheme=config.scheme,
        hyperviscosity_enabled=hyperviscosity_enabled_for_config(config),
        hyperviscosity_applications=hyperviscosity_applications,
        final_sensor_fraction=float(v[_lib.RDIAG_MASK] / n),
        density_min=float(v[_lib.RDIAG_RHO_MIN]), density_max=float(v[_lib.RDIAG_RHO_MAX]),
        pressure_min=float(v[_lib.RDIAG_P_MIN]), pressure_max=float(v[_lib.RDIAG_P_MAX]),
        max_abs_vorticity=float(v[_lib.RDIAG_VORT_MAX]),
        has_nonfinite=bool(v[_lib.RDIAG_NONFINITE_Q] > 0 or v[_lib.RDIAG_NONFINITE_RP] > 0),
        density_positive=bool(v[_lib.RDIAG_RHO_NONPOS] == 0),
        pressure_positive=bool(v[_lib.RDIAG_P_NONPOS] == 0),
        time=time, steps=steps)


def print_diagnostic_summary(diagnostics: RiemannDiagnostics) -> None:
    """riemann_config3.py:653-664."""
    d = diagnostics
    print("Diagnostic summary:")
    print(f"  backend              : {d.backend}")
    print(f"  scheme               : {d.scheme}")
    print(f"  density min/max      : {d.density_min:.8e} / {d.density_max:.8e}")
    print(f"  pressure min/max     : {d.pressure_min:.8e} / {d.pressure_max:.8e}")
    print(f"  max |vorticity|      : {d.max_abs_vorticity:.8e}")
    print(f"  NaN/Inf present      : {d.has_nonfinite}")
    print(f"  density positive     : {d.density_positive}")
    print(f"  pressure positive    : {d.pressure_positive}")
    print(f"  final sensor fraction: {d.final_sensor_fraction:.6f}")
    print(f"  hyperviscosity       : {d.hyperviscosity_enabled} ({d.hyperviscosity_applications} applications)")


class RiemannSolver:
    """The reference step loop (riemann_config3.py:681-749) as an object:
    ``step()`` runs one fused SSP-RK3 step, then the scheduled hyperviscosity
    pass with a fresh mask and the guard."""

    def __init__(self, config: RiemannConfig3, q0=None):
        if config.backend not in ("hip", "auto"):
            raise ValueError("this package implements the 'hip' backend")
        self.config = config
        eq = config.equation
        self.equation = eq
        dt_ = config.torch_dtype
        q = config.initial_state() if q0 is None else q0
        self.sensor = NonPeriodicEulerShockSensor2D(
            config.domain, eq, width=config.sensor_width, compression_threshold=config.compression_threshold,
            jump_threshold=config.jump_threshold, shear_threshold=config.shear_threshold,
            boundary_guard=config.boundary_guard)
        self.operator = NonPeriodicHybridEuler2DOperator(config.domain, eq, self.sensor, scheme=config.scheme)
        self.rhs = self.operator.native_rhs(dt_, config.guard_cells)
        q = self.rhs.clean(q)  # outflow guard of enforce(q), riemann_config3.py:685-686
        validate_raw_physical_state(q, eq, "initial state")
        self.dt, self.n_steps = self.operator.fixed_time_step(q, config.cfl, config.tfinal)
        self.hv_enabled = hyperviscosity_enabled_for_config(config)
        self.hv = LocalHyperviscosity2D(
            config.domain, mn=config.mn, density_weight=config.hyperviscosity_density_weight,
            momentum_weight=config.hyperviscosity_momentum_weight,
            energy_weight=config.hyperviscosity_energy_weight) if self.hv_enabled else None
        self.integrator = SSPRK3()
        self.q = q
        self.steps = 0
        self.hv_applications = 0
        self.time = 0.0
        self._spare = torch.empty_like(q)
        if self.hv_enabled:
            self._hv_plan = get_plan(self.hv.spec(dt_, eq, config.guard_cells))
            self._sensor_plan = get_plan(self.sensor.plan_spec(dt_))
            self._mask = torch.empty(q.shape[1:], dtype=torch.uint8, device=q.device)

    def step(self, validate: bool = True):
        c = self.config
        q = self.integrator.step(self.q, self.rhs, self.dt, clean=self.rhs.clean, out=self._spare)
        self._spare = self.q
        if self.hv_enabled and (self.steps + 1) % c.hyperviscosity_interval == 0:
            weno = self._sensor_plan.sensor(q, enforce_first=False, out=self._mask)
            q_new = self._hv_plan.hyperviscosity(q, weno, apply_guard=c.guard_cells > 0, out=self._spare,
                                                 invert=True)
            self._spare = q
            q = q_new
            self.hv_applications += 1
        self.steps += 1
        self.time = self.steps * self.dt
        if validate and (self.steps % c.progress_every == 0 or self.steps == self.n_steps):
            validate_raw_physical_state(q, self.equation, f"state at step {self.steps}")
        self.q = q
        return q

    def capture(self, before_capture=None):
        """Record one step (3 fused stages + the hyperviscosity pass) as a HIP
        graph.  The fixed dt and the buffer rotation make every step launch the
        same kernels on the same buffers when hyperviscosity runs every step."""
        c = self.config
        if self.hv_enabled and c.hyperviscosity_interval != 1:
            raise ValueError("graph capture needs hyperviscosity every step (or none)")
        if not self.hv_enabled:
            raise ValueError("graph capture is implemented for the hyperviscosity-every-step loop")
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._step_kernels()  # warm-up on the capture stream (plans, buffers)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        if before_capture is not None:
            before_capture()
        self._graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._graph):
            self._step_kernels()
        torch.cuda.synchronize()
        return self._graph

    def _step_kernels(self):
        """Step body without host-side bookkeeping (used for capture); the
        buffers come back to the same places after the hyperviscosity pass."""
        q = self.integrator.step(self.q, self.rhs, self.dt, clean=self.rhs.clean, out=self._spare)
        weno = self._sensor_plan.sensor(q, enforce_first=False, out=self._mask)
        self._hv_plan.hyperviscosity(q, weno, apply_guard=self.config.guard_cells > 0, out=self.q, invert=True)

    def replay(self, n: int = 1):
        """Advance n steps by replaying the captured graph."""
        for _ in range(n):
            self._graph.replay()
            self.steps += 1
            self.hv_applications += 1
        self.time = self.steps * self.dt

    def run(self, max_steps: int | None = None):
        n = self.n_steps if max_steps is None else min(max_steps, self.n_steps)
        while self.steps < n:
            self.step()
        return self.q

    def diagnostics(self) -> RiemannDiagnostics:
        return _diagnostics(self.q, self.config, self.time, self.steps, self.hv_applications, self.sensor)


def run_riemann(config: RiemannConfig3):
    """riemann_config3.py:681-749: returns (q, RiemannDiagnostics), with the
    reference's progress lines and summary."""
    print("Using array backend: hip")
    s = RiemannSolver(config)
    label = getattr(config, "configuration_label", f"Config {config.configuration_number}")
    print(f"Riemann {label}: nx={config.nx}, ny={config.ny}, dt={s.dt:.5e}, steps={s.n_steps}, "
          f"tfinal={config.tfinal}")
    note = "enabled on compact-FD nodes only" if s.hv_enabled else "disabled"
    if config.scheme == "weno" and config.mn > 0.0:
        note += " (WENO-only scheme)"
    print(f"scheme={config.scheme}; local hyperviscosity {note} (mn={config.mn})")
    while s.steps < s.n_steps:
        s.step()
        if s.steps % config.progress_every == 0 or s.steps == s.n_steps:
            print(f"step={s.steps:6d}, time={s.time:.6f}")
    d = s.diagnostics()
    validate_raw_physical_state(s.q, config.equation, "final state")
    print_diagnostic_summary(d)
    return s.q, d


def run_riemann_config3(config: RiemannConfig3):
    return run_riemann(config)


def make_run_id(prefix: str) -> str:
    """Timestamped run identifier (OOP/run_utils.py:14-18)."""
    import uuid
    from datetime import datetime

    return f"{prefix}_{datetime.now().strftime('%Y%m%d_%H%M%S_%f')}_{uuid.uuid4().hex[:8]}"


def save_riemann_npz(path: Path, q, config: RiemannConfig3, diagnostics: RiemannDiagnostics) -> Path:
    """riemann_config3.py:497-519: grid, state, primitive planes, omega_z,
    time, steps and the configuration and diagnostics as JSON strings."""
    d = config.domain
    _v, planes = _check_plan(q, config.equation, d).riemann_diagnostics(q, fields=True)
    f = planes.cpu().numpy()
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(
        path, x=d.x, y=d.y, q=q.double().cpu().numpy(), rho=f[0], u=f[1], v=f[2], pressure=f[3], omega_z=f[4],
        time=np.asarray(diagnostics.time), steps=np.asarray(diagnostics.steps),
        configuration_number=np.asarray(config.configuration_number),
        config_json=np.asarray(json.dumps(riemann_config_dict(config), default=str)),
        diagnostics_json=np.asarray(json.dumps(asdict(diagnostics), default=str)),
    )
    return path


def load_riemann_npz(path: Path):
    """riemann_config3.py:522-532: (npz data, RiemannConfig3 of its config_json)."""
    data = np.load(path, allow_pickle=False)
    conf = json.loads(str(data["config_json"])) if "config_json" in data else {}
    fields = RiemannConfig3.__dataclass_fields__
    return data, RiemannConfig3(**{k: v for k, v in conf.items() if k != "configuration_number" and k in fields})


def save_run_outputs(q, config: RiemannConfig3, diagnostics: RiemannDiagnostics, output_dir: Path,
                     run_id: str | None = None, plot: bool = True, **plot_options) -> dict:
    """riemann_config3.py:756-810: <output_dir>/<run_id>/config.json,
    diagnostics.json and <run_id>_final.npz.  Figures are not drawn by this
    backend: the npz is the input of the reference's plot_from_npz."""
    run_id = run_id or make_run_id(f"riemann{config.configuration_number}")
    run_dir = Path(output_dir) / run_id
    run_dir.mkdir(parents=True, exist_ok=True)
    for name, payload in (("config.json", riemann_config_dict(config)), ("diagnostics.json", asdict(diagnostics))):
        (run_dir / name).write_text(json.dumps(payload, indent=2, sort_keys=True))
    npz_path = save_riemann_npz(run_dir / f"{run_id}_final.npz", q, config, diagnostics)
    if plot:
        print(f"figure not drawn (plot_from_npz of the reference reads {npz_path.name})")
    return {"npz": npz_path}


def riemann_config_dict(config: RiemannConfig3) -> dict:
    values = {"configuration_number": config.configuration_number, **asdict(config)}
    for k in ("scenario_name", "configuration_label"):
        v = getattr(config, k, None)
        if v is not None:
            values[k] = v
    return values
