"""Periodic (HIT) hybrid compact/WENO7 operator, FP64 NumPy.

Restates reference OOP/spatial_operator.py:87-124 (weno7_flux), :147-211
(jump indicator, dilation, smooth and WENO interface fluxes), :227-249 (FFT
circulant compact solve), :252-377 (sensor, operator, dt) and :380-435
(hyperviscosity).  Interface index i is x_{i+1/2}; arrays are [.., ny, nx]
with x on the last axis.  ``sh(a, k)[i] == a[i + k]`` (periodic).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .equations import Gas, enforce, flux, primitives

ALPHA_C = 3.0 / 8.0  # Lele alpha, OOP/spatial_operator.py:240


def sh(a, k, axis=-1):
    """Periodic shift: sh(a, k)[i] = a[i + k] along ``axis``."""
    return np.roll(a, -k, axis=axis)


def weno7_parts(v1, v2, v3, v4, v5, v6, v7):
    """Candidate values q0..q3 and smoothness indicators b0..b3 of WENO7-JS,
    OOP/spatial_operator.py:89-117.  b0 carries the reference's coefficient
    544 on v1^2 (:95; the published Balsara-Shu value is 547)."""
    q0 = -(1 / 4) * v1 + (13 / 12) * v2 - (23 / 12) * v3 + (25 / 12) * v4
    q1 = (1 / 12) * v2 - (5 / 12) * v3 + (13 / 12) * v4 + (1 / 4) * v5
    q2 = -(1 / 12) * v3 + (7 / 12) * v4 + (7 / 12) * v5 - (1 / 12) * v6
    q3 = (1 / 4) * v4 + (13 / 12) * v5 - (5 / 12) * v6 + (1 / 12) * v7
    b0 = (v1 * (544 * v1 - 3882 * v2 + 4642 * v3 - 1854 * v4)
          + v2 * (7043 * v2 - 17246 * v3 + 7042 * v4)
          + v3 * (11003 * v3 - 9402 * v4)
          + 2107 * v4**2)
    b1 = (v2 * (267 * v2 - 1642 * v3 + 1602 * v4 - 494 * v5)
          + v3 * (2843 * v3 - 5966 * v4 + 1922 * v5)
          + v4 * (3443 * v4 - 2522 * v5)
          + 547 * v5**2)
    b2 = (v3 * (547 * v3 - 2522 * v4 + 1922 * v5 - 494 * v6)
          + v4 * (3443 * v4 - 5966 * v5 + 1602 * v6)
          + v5 * (2843 * v5 - 1642 * v6)
          + 267 * v6**2)
    b3 = (v4 * (2107 * v4 - 9402 * v5 + 7042 * v6 - 1854 * v7)
          + v5 * (11003 * v5 - 17246 * v6 + 4642 * v7)
          + v6 * (7043 * v6 - 3882 * v7)
          + 547 * v7**2)
    return (q0, q1, q2, q3), (b0, b1, b2, b3)


def weno7(v1, v2, v3, v4, v5, v6, v7):
    """WENO7-JS reconstruction, OOP/spatial_operator.py:87-124: eps 1e-10,
    p = 2, linear weights (1, 12, 18, 4)/35."""
    eps = 1e-10
    (q0, q1, q2, q3), (b0, b1, b2, b3) = weno7_parts(v1, v2, v3, v4, v5, v6, v7)
    a0 = (1 / 35) / (eps + b0) ** 2
    a1 = (12 / 35) / (eps + b1) ** 2
    a2 = (18 / 35) / (eps + b2) ** 2
    a3 = (4 / 35) / (eps + b3) ** 2
    asum = a0 + a1 + a2 + a3
    return (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) / asum


def jump_indicator(f, axis):
    """OOP/spatial_operator.py:147-166: max of left/right relative jumps and
    relative curvature, eps = 1e-14 + 1e-12 * global max|f|."""
    eps = 1e-14 + 1e-12 * np.max(np.abs(f))
    fr = sh(f, 1, axis)
    fl = sh(f, -1, axis)
    jr = np.abs(fr - f)
    jr /= np.abs(fr) + np.abs(f) + eps
    jl = np.abs(f - fl)
    jl /= np.abs(f) + np.abs(fl) + eps
    cv = np.abs(fr - 2.0 * f + fl)
    cv /= np.abs(fr) + 2.0 * np.abs(f) + np.abs(fl) + eps
    return np.maximum(np.maximum(jl, jr), cv)


def dilate(mask, width):
    """OOP/spatial_operator.py:169-176: periodic (2w+1)^2 box OR, done as the
    reference does it (one double shift per offset)."""
    out = np.array(mask, copy=True)
    for oy in range(-width, width + 1):
        for ox in range(-width, width + 1):
            if ox or oy:
                out |= np.roll(np.roll(mask, oy, axis=-2), ox, axis=-1)
    return out


@dataclass(frozen=True)
class SensorParams:
    width: int = 4
    compression_threshold: float = 2.5
    jump_threshold: float = 0.04
    shear_threshold: float | None = None


def sensor(q, gas: Gas, dx, dy, sp: SensorParams):
    """OOP/spatial_operator.py:263-304: compression | shear | jumps(rho, p, e),
    then periodic dilation.  Returns bool[ny, nx] (True = WENO node)."""
    rho, u, v, p = primitives(q, gas)
    ux = (sh(u, 1) - sh(u, -1)) / (2.0 * dx)
    uy = (sh(u, 1, -2) - sh(u, -1, -2)) / (2.0 * dy)
    vx = (sh(v, 1) - sh(v, -1)) / (2.0 * dx)
    vy = (sh(v, 1, -2) - sh(v, -1, -2)) / (2.0 * dy)
    div = ux + vy
    div_rms = float(np.sqrt(np.mean(div**2)))
    comp = np.zeros_like(div, dtype=bool)
    if div_rms > 1e-12:
        comp = div < -sp.compression_threshold * div_rms
    shear = np.zeros_like(div, dtype=bool)
    if sp.shear_threshold is not None:
        w = vx - uy
        w_rms = float(np.sqrt(np.mean(w**2)))
        if w_rms > 1e-12:
            shear = np.abs(w) > sp.shear_threshold * w_rms
    e = p / (rho * (gas.gamma - 1.0))
    jr = np.maximum(jump_indicator(rho, -1), jump_indicator(rho, -2))
    jp = np.maximum(jump_indicator(p, -1), jump_indicator(p, -2))
    je = np.maximum(jump_indicator(e, -1), jump_indicator(e, -2))
    raw = comp | shear | (jr > sp.jump_threshold) | (jp > sp.jump_threshold) | (je > sp.jump_threshold)
    return dilate(raw, sp.width)


def smooth_flux(pf):
    """OOP/spatial_operator.py:179-189: explicit interface flux whose
    difference is Lele's 8th-order (alpha 3/8) compact right-hand side."""
    a, b, c = 25 / 32, 1 / 20, -1 / 480
    return (c * (sh(pf, 3) + sh(pf, -2))
            + (b + c) * (sh(pf, 2) + sh(pf, -1))
            + (a + b + c) * (sh(pf, 1) + pf))


def weno_flux(q, pf, alpha):
    """OOP/spatial_operator.py:192-211: global Lax-Friedrichs split, WENO7 of
    f+ on i-3..i+3 and of f- on i+4..i-2."""
    fp = 0.5 * (pf + alpha * q)
    fm = 0.5 * (pf - alpha * q)
    return (weno7(sh(fp, -3), sh(fp, -2), sh(fp, -1), fp, sh(fp, 1), sh(fp, 2), sh(fp, 3))
            + weno7(sh(fm, 4), sh(fm, 3), sh(fm, 2), sh(fm, 1), fm, sh(fm, -1), sh(fm, -2)))


def compact_solve(F, h, alpha=ALPHA_C):
    """OOP/spatial_operator.py:246-249: raw = (F_i - F_{i-1})/h, then the
    cyclic (alpha, 1, alpha) system solved by FFT along the last axis."""
    n = F.shape[-1]
    eig = 1.0 + 2.0 * alpha * np.cos(2.0 * np.pi * np.fft.fftfreq(n))
    raw = (F - sh(F, -1)) / h
    return np.fft.ifft(np.fft.fft(raw, axis=-1) / eig, axis=-1).real


def axis_derivative(q_ax, pf_ax, mask_ax, h, gas: Gas, normal: int):
    """OOP/spatial_operator.py:321-342 for lines along the last axis."""
    rho, u, v, p = primitives(q_ax, gas)
    un = u if normal == 1 else v
    c = np.sqrt(gas.gamma * p / rho)
    alpha = float(np.max(np.abs(un) + c))
    smooth = smooth_flux(pf_ax)
    raw = weno_flux(q_ax, pf_ax, alpha)
    wf = ALPHA_C * sh(raw, 1)
    wf += raw
    wf += ALPHA_C * sh(raw, -1)
    iface = mask_ax | sh(mask_ax, 1)
    hybrid = np.where(iface[None, ...], wf, smooth)
    return compact_solve(hybrid, h)


def rhs(q, gas: Gas, dx, dy, sp: SensorParams):
    """OOP/spatial_operator.py:344-367: enforce -> sensor -> d/dx -> d/dy on the
    transposed view -> -dFx - dFy."""
    qs = enforce(q, gas)
    m = sensor(qs, gas, dx, dy, sp)
    d_x = axis_derivative(qs, flux(qs, gas, "x"), m, dx, gas, 1)
    qt = np.moveaxis(qs, -2, -1)
    ft = np.moveaxis(flux(qs, gas, "y"), -2, -1)
    mt = np.moveaxis(m, -2, -1)
    d_y = np.moveaxis(axis_derivative(qt, ft, mt, dy, gas, 2), -1, -2)
    return -d_x - d_y


def stable_dt(q, gas: Gas, dx, dy, cfl, remaining=None):
    """OOP/spatial_operator.py:369-377."""
    rho, u, v, p = primitives(q, gas)
    c = np.sqrt(gas.gamma * p / rho)
    sr = (np.abs(u) + c) / dx
    sr += (np.abs(v) + c) / dy
    dt = cfl / float(np.max(sr))
    if remaining is not None:
        dt = min(dt, remaining)
    return dt


def _lap(a, dx, dy):
    lx = (sh(a, 1) - 2.0 * a + sh(a, -1)) / dx**2
    ly = (sh(a, 1, -2) - 2.0 * a + sh(a, -1, -2)) / dy**2
    return lx + ly


def hyperviscosity(q, gas: Gas, dx, dy, mn, active_mask, weights=(0.25, 1.0, 1.0, 0.25)):
    """OOP/spatial_operator.py:407-435: q - mn*w*h^4*active*lap(lap(q)), then
    enforce.  ``active_mask`` (compact nodes) is required."""
    if active_mask is None:
        raise ValueError("active_mask is required: hyperviscosity may only be applied on compact-FD nodes")
    if active_mask.shape != q.shape[-2:]:
        raise ValueError(f"active_mask shape {active_mask.shape} does not match grid {q.shape[-2:]}")
    h = min(dx, dy)
    bih = _lap(_lap(q, dx, dy), dx, dy)
    w = np.array([weights[0], weights[1], weights[2], weights[3]], dtype=q.dtype)[:, None, None]
    filtered = q - mn * w * h**4 * active_mask[None, :, :] * bih
    return enforce(filtered, gas)
