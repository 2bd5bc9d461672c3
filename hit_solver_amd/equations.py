"""Gas algebra on device tensors, API of reference OOP/parallel/equations.py.

``enforce_physical_state``, ``primitive_from_conservative`` and
``viscous_rhs`` run as HIP kernels; the remaining helpers are small torch
expressions used outside the hot path (setup, diagnostics).
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from .runtime import PlanSpec, get_plan


def _shape_of(q):
    if q.dim() != 3 or q.shape[0] != 4:
        raise ValueError(f"state must have shape (4, ny, nx), got {tuple(q.shape)}")
    return int(q.shape[2]), int(q.shape[1])


@dataclass(frozen=True)
class EulerEquation2D:
    """OOP/equations.py:125-208 on the hip backend."""

    gamma: float = 1.4
    rho_floor: float = 1e-10
    pressure_floor: float = 1e-10

    def _plan(self, q, **extra):
        nx, ny = _shape_of(q)
        return get_plan(PlanSpec(nx=nx, ny=ny, dx=1.0, dy=1.0, periodic=True, dtype=q.dtype, gamma=self.gamma,
                                 rho_floor=self.rho_floor, p_floor=self.pressure_floor, **extra))

    def primitive_from_conservative(self, q):
        return self._plan(q).primitives(q)

    def conservative_from_primitive(self, rho, u, v, pressure):
        rho, u, v, pressure = torch.broadcast_tensors(rho, u, v, pressure)
        rho = torch.clamp_min(rho, self.rho_floor)
        pressure = torch.clamp_min(pressure, self.pressure_floor)
        q = torch.zeros((4, *rho.shape), dtype=rho.dtype, device=rho.device)
        q[0] = rho
        q[1] = rho * u
        q[2] = rho * v
        q[3] = pressure / (self.gamma - 1.0) + 0.5 * rho * (u**2 + v**2)
        return q

    def enforce_physical_state(self, q):
        return self._plan(q).enforce(q, apply_guard=False)

    def flux_x(self, q):
        rho, u, v, p = self.primitive_from_conservative(q)
        return torch.stack([q[1], rho * u**2 + p, rho * u * v, (q[3] + p) * u])

    def flux_y(self, q):
        rho, u, v, p = self.primitive_from_conservative(q)
        return torch.stack([q[2], rho * u * v, rho * v**2 + p, (q[3] + p) * v])

    def sound_speed(self, q):
        rho, _u, _v, p = self.primitive_from_conservative(q)
        return torch.sqrt(self.gamma * p / rho)

    def max_wave_speed(self, q) -> float:
        rho, u, v, p = self.primitive_from_conservative(q)
        c = torch.sqrt(self.gamma * p / rho)
        return float(torch.max(torch.maximum(u.abs() + c, v.abs() + c)))


@dataclass(frozen=True)
class CompressibleNavierStokes2D(EulerEquation2D):
    """OOP/equations.py:211-267: viscous and heat-conduction terms (periodic)."""

    viscosity: float = 0.0
    prandtl: float = 0.72
    gas_constant: float = 1.0

    def temperature(self, rho, pressure):
        return pressure / (rho * self.gas_constant)

    def viscous_rhs(self, q, dx: float, dy: float):
        if self.viscosity == 0.0:
            return torch.zeros_like(q)
        nx, ny = _shape_of(q)
        plan = get_plan(PlanSpec(nx=nx, ny=ny, dx=dx, dy=dy, periodic=True, dtype=q.dtype, gamma=self.gamma,
                                 rho_floor=self.rho_floor, p_floor=self.pressure_floor, viscosity=self.viscosity,
                                 prandtl=self.prandtl, gas_constant=self.gas_constant))
        return plan.viscous_rhs(q)
