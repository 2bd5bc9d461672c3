"""Ideal-gas algebra in conservative variables q = [rho, rho u, rho v, E].

Restates reference OOP/equations.py:125-267 (and its backend-aware twin
OOP/parallel/equations.py:10-141) in the same floating-point operation order.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Gas:
    gamma: float = 1.4
    rho_floor: float = 1e-10
    p_floor: float = 1e-10
    viscosity: float = 0.0
    prandtl: float = 0.72
    gas_constant: float = 1.0


def primitives(q, gas: Gas):
    """OOP/equations.py:136-145: floored rho and p; u, v = q1/rho, q2/rho."""
    rho = np.maximum(q[0], gas.rho_floor)
    u = q[1] / rho
    v = q[2] / rho
    ke = 0.5 * rho * (u**2 + v**2)
    p = (gas.gamma - 1.0) * (q[3] - ke)
    return rho, u, v, np.maximum(p, gas.p_floor)


def conservative(rho, u, v, p, gas: Gas):
    """OOP/equations.py:147-163."""
    rho, u, v, p = np.broadcast_arrays(rho, u, v, p)
    rho = np.maximum(rho, gas.rho_floor)
    p = np.maximum(p, gas.p_floor)
    q = np.zeros((4, *rho.shape), dtype=np.result_type(rho, u, v, p))
    q[0] = rho
    q[1] = rho * u
    q[2] = rho * v
    q[3] = p / (gas.gamma - 1.0) + 0.5 * rho * (u**2 + v**2)
    return q


def enforce(q, gas: Gas):
    """OOP/equations.py:165-181: floors rho, re-forms momentum as rho*(q/rho),
    resets E where the pressure is at or below the floor.  Returns a copy."""
    out = np.array(q, copy=True)
    rho = np.maximum(out[0], gas.rho_floor)
    u = out[1] / rho
    v = out[2] / rho
    ke = 0.5 * rho * (u**2 + v**2)
    p = (gas.gamma - 1.0) * (out[3] - ke)
    out[0] = rho
    out[1] = rho * u
    out[2] = rho * v
    out[3] = np.where(p > gas.p_floor, out[3], gas.p_floor / (gas.gamma - 1.0) + ke)
    return out


def flux(q, gas: Gas, axis: str):
    """OOP/equations.py:183-199 (flux_x for axis 'x', flux_y for 'y')."""
    rho, u, v, p = primitives(q, gas)
    f = np.zeros_like(q)
    if axis == "x":
        f[0] = q[1]
        f[1] = rho * u**2 + p
        f[2] = rho * u * v
        f[3] = (q[3] + p) * u
    else:
        f[0] = q[2]
        f[1] = rho * u * v
        f[2] = rho * v**2 + p
        f[3] = (q[3] + p) * v
    return f


def _dper(a, h, axis):
    # OOP/equations.py:229-230: second-order central difference, periodic
    return (np.roll(a, -1, axis=axis) - np.roll(a, 1, axis=axis)) / (2.0 * h)


def viscous_rhs(q, gas: Gas, dx: float, dy: float):
    """OOP/equations.py:232-267: Stokes stress + Fourier conduction, nested
    central differences, k = mu*cp/Pr.  Zero when the viscosity is zero."""
    if gas.viscosity == 0.0:
        return np.zeros_like(q)
    rho, u, v, p = primitives(q, gas)
    temp = p / (rho * gas.gas_constant)
    mu = gas.viscosity
    cp = gas.gamma * gas.gas_constant / (gas.gamma - 1.0)
    kcond = mu * cp / gas.prandtl

    ux = _dper(u, dx, -1)
    uy = _dper(u, dy, -2)
    vx = _dper(v, dx, -1)
    vy = _dper(v, dy, -2)
    div = ux + vy
    txx = 2.0 * mu * ux - (2.0 / 3.0) * mu * div
    tyy = 2.0 * mu * vy - (2.0 / 3.0) * mu * div
    txy = mu * (uy + vx)
    tx = _dper(temp, dx, -1)
    ty = _dper(temp, dy, -2)

    r = np.zeros_like(q)
    r[1] = _dper(txx, dx, -1)
    r[1] += _dper(txy, dy, -2)
    r[2] = _dper(txy, dx, -1)
    r[2] += _dper(tyy, dy, -2)
    ex = u * txx + v * txy + kcond * tx
    ey = u * txy + v * tyy + kcond * ty
    r[3] = _dper(ex, dx, -1)
    r[3] += _dper(ey, dy, -2)
    return r
