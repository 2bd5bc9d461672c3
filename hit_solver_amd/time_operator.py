"""SSP-RK3 and the minimal time loop, API of reference OOP/time_operator.py:17-69.

``SSPRK3.step`` accepts any callable RHS, as the reference does.  When the
RHS is a :class:`NativeRHS` (the operator's RHS plus its source closure, as
built by the drivers) and ``clean`` is that RHS's own ``clean``, each stage
runs as one fused launch sequence on the device (``hitx_rk3_stage``): sweeps,
sources, Shu-Osher combination and clean together.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable

import torch


class NativeRHS:
    """RHS callable backed by a stage plan.  ``fx``/``fy`` are the forcing
    accelerations held fixed over the step (OOP/hit2d.py:851)."""

    def __init__(self, plan, fx=None, fy=None):
        self.plan = plan
        self.fx = fx
        self.fy = fy
        self._bufs = None

    def __call__(self, q):
        # the operator RHS alone (without the fused sources); the stage path
        # is what the integrator uses
        return self.plan.rhs(q)

    def clean(self, q):
        """enforce_physical_state, then the outflow guard when the plan has one."""
        return self.plan.enforce(q, apply_guard=True)

    def buffers(self, q):
        if self._bufs is None or self._bufs[0].shape != q.shape or self._bufs[0].dtype != q.dtype:
            self._bufs = (torch.empty_like(q), torch.empty_like(q))
        return self._bufs


State = torch.Tensor
RHS = Callable[[State], State]


@dataclass(frozen=True)
class SSPRK3:
    """Shu-Osher SSP-RK3 (OOP/time_operator.py:17-26)."""

    def step(self, state, rhs, dt: float, clean=None, out=None):
        if isinstance(rhs, NativeRHS) and (clean is None or getattr(clean, "__self__", None) is rhs):
            q1, q2 = rhs.buffers(state)
            out = torch.empty_like(state) if out is None else out
            p = rhs.plan
            p.stage(0, state, state, q1, dt, rhs.fx, rhs.fy)
            p.stage(1, state, q1, q2, dt, rhs.fx, rhs.fy)
            p.stage(2, state, q2, out, dt, rhs.fx, rhs.fy)
            return out
        c = clean if clean is not None else (lambda v: v)
        q0 = state
        q1 = c(q0 + dt * rhs(q0))
        q2 = c(0.75 * q0 + 0.25 * (q1 + dt * rhs(q1)))
        return c((1.0 / 3.0) * q0 + (2.0 / 3.0) * (q2 + dt * rhs(q2)))


class PeriodicBoundary:
    """OOP/boundary_conditions.py:17-24: a copy (stencils wrap by index)."""

    def apply(self, state):
        return state.clone()


@dataclass
class TimeOperator:
    """OOP/time_operator.py:29-69."""

    state: torch.Tensor
    rhs: RHS
    dt: float
    boundary: PeriodicBoundary = field(default_factory=PeriodicBoundary)
    integrator: SSPRK3 = field(default_factory=SSPRK3)
    state_filter: Callable | None = None
    step_hooks: list = field(default_factory=list)

    def __post_init__(self):
        self.state = self.clean(self.state)

    def clean(self, state):
        if self.state_filter is not None:
            state = self.state_filter(state)
        return self.boundary.apply(state)

    def step(self, step_index: int):
        self.state = self.integrator.step(self.state, self.rhs, self.dt, clean=self.clean)
        for hook in self.step_hooks:
            self.state = self.clean(hook(step_index, self.state))
        return self.state

    def run(self, t_end: float, save_every: int | None = None):
        import math

        n_steps = int(math.ceil(t_end / self.dt))
        if n_steps > 0:
            self.dt = t_end / n_steps
        if save_every is None:
            save_every = n_steps
        times = [0.0]
        states = [self.state.clone()]
        for i in range(n_steps):
            self.step(i)
            if (i + 1) % save_every == 0 or i + 1 == n_steps:
                times.append((i + 1) * self.dt)
                states.append(self.state.clone())
        return torch.tensor(times, dtype=torch.float64), states
