"""Periodic HIT operator classes on the hip backend.

Same class names, constructor arguments and methods as the reference's
OOP/spatial_operator.py:227-435 (and the backend-aware twins in
OOP/parallel/spatial_operator.py).  Every method launches HIP kernels through
the hitx C ABI; there is no host fallback.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import torch

from .domain import Domain2D
from .equations import CompressibleNavierStokes2D, EulerEquation2D
from .runtime import PlanSpec, get_plan


def _grid(q, domain: Domain2D):
    if q.dim() != 3 or q.shape[0] != 4 or q.shape[1] != domain.ny or q.shape[2] != domain.nx:
        raise ValueError(f"state shape {tuple(q.shape)} does not match (4, {domain.ny}, {domain.nx})")


@dataclass
class PeriodicLineCompactDerivative:
    """Holds the compact-scheme constants (OOP/spatial_operator.py:227-249).  The
    periodic (alpha, 1, alpha) solve itself runs inside the sweep kernels."""

    n: int
    dx: float
    alpha: float = 3.0 / 8.0


@dataclass(frozen=True)
class PeriodicEulerShockSensor2D:
    """OOP/spatial_operator.py:252-304."""

    domain: Domain2D
    equation: EulerEquation2D
    width: int = 4
    compression_threshold: float = 2.5
    jump_threshold: float = 0.04
    shear_threshold: float | None = None

    def plan_spec(self, dtype, **extra) -> PlanSpec:
        eq = self.equation
        return PlanSpec(
            nx=self.domain.nx, ny=self.domain.ny, dx=self.domain.dx, dy=self.domain.dy, periodic=True,
            dtype=dtype, gamma=eq.gamma, rho_floor=eq.rho_floor, p_floor=eq.pressure_floor,
            sensor_width=self.width, compression_threshold=self.compression_threshold,
            jump_threshold=self.jump_threshold, shear_threshold=self.shear_threshold, **extra,
        )

    def detect_u8(self, q):
        _grid(q, self.domain)
        return get_plan(self.plan_spec(q.dtype)).sensor(q, enforce_first=False)

    def detect(self, q):
        """bool[ny, nx], True at WENO nodes."""
        return self.detect_u8(q).bool()


@dataclass
class PeriodicHybridEuler2DOperator:
    """OOP/spatial_operator.py:307-377: hybrid compact/WENO7 RHS."""

    domain: Domain2D
    equation: EulerEquation2D
    sensor: PeriodicEulerShockSensor2D
    compact_x: PeriodicLineCompactDerivative = field(init=False)
    compact_y: PeriodicLineCompactDerivative = field(init=False)

    def __post_init__(self):
        self.compact_x = PeriodicLineCompactDerivative(self.domain.nx, self.domain.dx)
        self.compact_y = PeriodicLineCompactDerivative(self.domain.ny, self.domain.dy)

    def plan(self, dtype, **extra):
        return get_plan(self.sensor.plan_spec(dtype, **extra))

    def rhs(self, q):
        _grid(q, self.domain)
        return self.plan(q.dtype).rhs(q)

    def stable_time_step(self, q, cfl: float, remaining_time: float | None = None) -> float:
        _grid(q, self.domain)
        max_sr, _, _ = self.plan(q.dtype).step_reduce(q)
        dt = cfl / max_sr
        if remaining_time is not None:
            dt = min(dt, remaining_time)
        return dt

    def native_rhs(self, dtype, **sources):
        """A stage-fused RHS for :class:`hit_solver_amd.time_operator.SSPRK3`:
        the operator RHS plus the sources given (see ``PlanSpec``)."""
        from .time_operator import NativeRHS

        return NativeRHS(self.plan(dtype, **sources))


@dataclass(frozen=True)
class PeriodicHyperviscosity2D:
    """OOP/spatial_operator.py:380-435: biharmonic filter on compact nodes."""

    domain: Domain2D
    mn: float = 0.002
    density_weight: float = 0.25
    momentum_weight: float = 1.0
    energy_weight: float = 0.25

    def coefficients(self):
        h = min(self.domain.dx, self.domain.dy)
        w = (self.density_weight, self.momentum_weight, self.momentum_weight, self.energy_weight)
        return tuple((self.mn * wk) * h**4 for wk in w)

    def apply(self, q, equation: EulerEquation2D, active_mask=None):
        if active_mask is None:
            raise ValueError(
                "active_mask is required: hyperviscosity may only be applied on compact-FD nodes"
            )
        if tuple(active_mask.shape) != tuple(q.shape[-2:]):
            raise ValueError(f"active_mask shape {tuple(active_mask.shape)} does not match grid {tuple(q.shape[-2:])}")
        _grid(q, self.domain)
        spec = PlanSpec(nx=self.domain.nx, ny=self.domain.ny, dx=self.domain.dx, dy=self.domain.dy, periodic=True,
                        dtype=q.dtype, gamma=equation.gamma, rho_floor=equation.rho_floor,
                        p_floor=equation.pressure_floor, hv_coef=self.coefficients())
        act = active_mask if active_mask.dtype == torch.uint8 else active_mask.to(torch.uint8)
        return get_plan(spec).hyperviscosity(q, act.contiguous(), apply_guard=False)


__all__ = [
    "CompressibleNavierStokes2D",
    "EulerEquation2D",
    "PeriodicEulerShockSensor2D",
    "PeriodicHybridEuler2DOperator",
    "PeriodicHyperviscosity2D",
    "PeriodicLineCompactDerivative",
]
