"""Lax-Liu Riemann configurations 3, 3 @ (0.8, 0.8) and 6, and the fixed-dt loop.

Restates reference OOP/problems/riemann_config3.py:19-99 (config 3 and its
initial state), :432-463 and :629-650 (raw density and pressure, vorticity,
run diagnostics), :497-519 (the saved planes), :667-749 (run_riemann),
OOP/problems/riemann_config3_08.py:9-19 and OOP/problems/riemann_config6.py:40-87.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import nonperiodic as N
from .equations import Gas, conservative, enforce, primitives
from .rk import ssprk3

# (rho, u, v, p) per quadrant: I = top-right, II = top-left, III = bottom-left, IV = bottom-right
QUADRANTS = {
    3: ((1.5, 0.0, 0.0, 1.5), (0.5323, 1.206, 0.0, 0.3), (0.138, 1.206, 1.206, 0.029),
        (0.5323, 0.0, 1.206, 0.3)),
    6: ((1.0, 0.75, -0.50, 1.0), (2.0, 0.75, 0.50, 1.0), (1.0, -0.75, 0.50, 1.0),
        (3.0, -0.75, -0.50, 1.0)),
}


@dataclass(frozen=True)
class RiemannParams:
    configuration: int = 3
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
    guard_cells: int = 2
    progress_every: int = 500

    @property
    def dx(self):
        return (self.x_max - self.x_min) / self.nx

    @property
    def dy(self):
        return (self.y_max - self.y_min) / self.ny

    def gas(self):
        return Gas(gamma=self.gamma)

    def sensor_params(self):
        return N.SensorParams(self.sensor_width, self.compression_threshold, self.jump_threshold,
                              self.shear_threshold, self.boundary_guard)

    def hv_weights(self):
        return (self.hyperviscosity_density_weight, self.hyperviscosity_momentum_weight,
                self.hyperviscosity_momentum_weight, self.hyperviscosity_energy_weight)


def config3(**kw):
    return RiemannParams(configuration=3, **kw)


def config3_08(**kw):
    """OOP/problems/riemann_config3_08.py:9-19."""
    kw.setdefault("x_split", 0.8)
    kw.setdefault("y_split", 0.8)
    kw.setdefault("hyperviscosity_interval", 5)
    return RiemannParams(configuration=3, **kw)


def config6(**kw):
    """OOP/problems/riemann_config6.py:40-47."""
    kw.setdefault("tfinal", 0.25)
    return RiemannParams(configuration=6, **kw)


def initial_state(p: RiemannParams):
    """Quadrant states on cell centres; top = y >= y_split, right = x >= x_split
    (riemann_config3.py:60-99, riemann_config6.py:49-87)."""
    x = p.x_min + p.dx * (np.arange(p.nx) + 0.5)
    y = p.y_min + p.dy * (np.arange(p.ny) + 0.5)
    X, Y = np.meshgrid(x, y, indexing="xy")
    top = Y >= p.y_split
    right = X >= p.x_split
    q1, q2, q3, q4 = QUADRANTS[p.configuration]
    prim = [np.full((p.ny, p.nx), q3[k]) for k in range(4)]
    for sel, st in ((top & right, q1), (top & ~right, q2), (~top & right, q4)):
        for k in range(4):
            prim[k][sel] = st[k]
    return conservative(*prim, p.gas())


def raw_density_pressure(q, gas: Gas):
    """riemann_config3.py:432-439 (no floors)."""
    rho = q[0]
    u = q[1] / rho
    v = q[2] / rho
    return rho, (gas.gamma - 1.0) * (q[3] - 0.5 * rho * (u**2 + v**2))


def vorticity_z(q, gas: Gas, dx, dy):
    """riemann_config3.py:459-463."""
    _rho, u, v, _p = primitives(q, gas)
    return N.fd(v, dx, -1) - N.fd(u, dy, -2)


def diagnostics(q, p: RiemannParams, time, steps, hv_applications):
    """riemann_config3.py:629-650 as a dict of the RiemannDiagnostics fields."""
    gas = p.gas()
    rho, pr = raw_density_pressure(q, gas)
    w = vorticity_z(q, gas, p.dx, p.dy)
    return dict(
        backend=None, scheme=p.scheme, hyperviscosity_enabled=hyperviscosity_enabled(p),
        hyperviscosity_applications=hv_applications,
        final_sensor_fraction=float(np.mean(N.sensor(q, gas, p.dx, p.dy, p.sensor_params()))),
        density_min=float(np.min(rho)), density_max=float(np.max(rho)), pressure_min=float(np.min(pr)),
        pressure_max=float(np.max(pr)), max_abs_vorticity=float(np.max(np.abs(w))),
        has_nonfinite=bool(np.any(~np.isfinite(q)) | np.any(~np.isfinite(rho)) | np.any(~np.isfinite(pr))),
        density_positive=bool(np.all(rho > 0.0)), pressure_positive=bool(np.all(pr > 0.0)), time=time,
        steps=steps)


def saved_fields(q, p: RiemannParams):
    """riemann_config3.py:497-519: the planes save_riemann_npz writes."""
    gas = p.gas()
    rho, u, v, pr = primitives(q, gas)
    return {"rho": rho, "u": u, "v": v, "pressure": pr, "omega_z": vorticity_z(q, gas, p.dx, p.dy)}


def hyperviscosity_enabled(p: RiemannParams):
    """riemann_config3.py:667-678."""
    return p.scheme == "hybrid" and p.mn > 0.0 and p.hyperviscosity_interval > 0


class RiemannRun:
    """riemann_config3.py:681-749 without plotting/output.  ``step()`` does one
    RK3 step plus the scheduled hyperviscosity pass."""

    def __init__(self, p: RiemannParams, q0=None):
        self.p = p
        self.gas = p.gas()
        q = initial_state(p) if q0 is None else q0
        q = N.outflow_guard(enforce(q, self.gas), p.guard_cells)
        N.validate(q, self.gas, "initial state")
        self.q = q
        self.sp = p.sensor_params()
        self.dt, self.n_steps = N.fixed_dt(q, self.gas, p.dx, p.dy, p.cfl, p.tfinal)
        self.hv = hyperviscosity_enabled(p)
        self.steps = 0
        self.hv_applications = 0
        self.time = 0.0

    def clean(self, s):
        return N.outflow_guard(enforce(s, self.gas), self.p.guard_cells)

    def rhs(self, s):
        return N.rhs(s, self.gas, self.p.dx, self.p.dy, self.sp, self.p.scheme)

    def step(self):
        p = self.p
        q = ssprk3(self.q, self.rhs, self.dt, clean=self.clean)
        if self.hv and (self.steps + 1) % p.hyperviscosity_interval == 0:
            m = N.sensor(q, self.gas, p.dx, p.dy, self.sp)
            q = N.hyperviscosity(q, ~m, self.gas, p.dx, p.dy, p.mn, p.hv_weights())
            q = N.outflow_guard(q, p.guard_cells)
            self.hv_applications += 1
        self.steps += 1
        self.time = self.steps * self.dt
        if self.steps % p.progress_every == 0 or self.steps == self.n_steps:
            N.validate(q, self.gas, f"state at step {self.steps}")
        self.q = q
        return q

    def run(self, max_steps=None):
        n = self.n_steps if max_steps is None else min(max_steps, self.n_steps)
        while self.steps < n:
            self.step()
        return self.q
