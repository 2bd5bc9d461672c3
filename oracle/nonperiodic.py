"""Non-periodic (Lax-Liu Riemann) hybrid operator, FP64 NumPy.

Restates reference OOP/problems/riemann_config3.py:121-429.  Interface index
j = 0..n is x_{j-1/2}; every stencil gather clamps to [0, n-1]; the compact
system has identity end rows and is solved by a row-loop Thomas algorithm.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .equations import Gas, enforce, flux, primitives
from .periodic import ALPHA_C, weno7


def gather(a, idx, axis=-1):
    """:121-125 take_clipped: edge-clamped gather."""
    n = a.shape[axis]
    return np.take(a, np.clip(np.asarray(idx), 0, n - 1).astype(np.int64), axis=axis)


def jump_indicator(f, axis):
    """:128-146: pair jumps credited to both nodes, clamped curvature."""
    m = np.moveaxis(f, axis, -1)
    eps = 1e-14 + 1e-12 * np.max(np.abs(m))
    pair = np.abs(m[..., 1:] - m[..., :-1])
    pair = pair / (np.abs(m[..., 1:]) + np.abs(m[..., :-1]) + eps)
    j = np.zeros_like(m)
    j[..., :-1] = np.maximum(j[..., :-1], pair)
    j[..., 1:] = np.maximum(j[..., 1:], pair)
    i = np.arange(m.shape[-1])
    left = gather(m, i - 1)
    right = gather(m, i + 1)
    cv = np.abs(right - 2.0 * m + left)
    cv = cv / (np.abs(right) + 2.0 * np.abs(m) + np.abs(left) + eps)
    return np.moveaxis(np.maximum(j, cv), -1, axis)


def dilate(mask, width):
    """:149-162: (2w+1)^2 box OR truncated at the domain edge."""
    out = np.array(mask, copy=True)
    ny, nx = mask.shape
    for oy in range(-width, width + 1):
        for ox in range(-width, width + 1):
            if ox == 0 and oy == 0:
                continue
            ys = slice(max(0, -oy), min(ny, ny - oy))
            yd = slice(max(0, oy), min(ny, ny + oy))
            xs = slice(max(0, -ox), min(nx, nx - ox))
            xd = slice(max(0, ox), min(nx, nx + ox))
            out[yd, xd] = out[yd, xd] | mask[ys, xs]
    return out


def iface_mask(mask):
    """:165-171: n+1 interface flags m[clip(j-1)] | m[clip(j)]."""
    j = np.arange(mask.shape[-1] + 1)
    return gather(mask, j - 1) | gather(mask, j)


def smooth_flux(pf):
    """:174-185: the periodic interface formula with clamped gathers."""
    a, b, c = 25 / 32, 1 / 20, -1 / 480
    i = np.arange(-1, pf.shape[-1])
    g = lambda k: gather(pf, i + k)
    return c * (g(3) + g(-2)) + (b + c) * (g(2) + g(-1)) + (a + b + c) * (g(1) + g(0))


def weno_flux(q, pf, alpha):
    """:188-212."""
    i = np.arange(-1, q.shape[-1])
    fp = 0.5 * (pf + alpha * q)
    fm = 0.5 * (pf - alpha * q)
    gp = lambda k: gather(fp, i + k)
    gm = lambda k: gather(fm, i + k)
    return (weno7(gp(-3), gp(-2), gp(-1), gp(0), gp(1), gp(2), gp(3))
            + weno7(gm(4), gm(3), gm(2), gm(1), gm(0), gm(-1), gm(-2)))


def compactify(w, alpha=ALPHA_C):
    """:215-222."""
    j = np.arange(w.shape[-1])
    return alpha * gather(w, j + 1) + w + alpha * gather(w, j - 1)


def compact_solve(F, h, alpha=ALPHA_C):
    """:237-265: raw = (F_{j+1} - F_j)/h; Thomas with (alpha, 1, alpha)
    interior rows and identity end rows, looping over rows as the reference."""
    raw = (F[..., 1:] - F[..., :-1]) / h
    lead = raw.shape[:-1]
    n = raw.shape[-1]
    r = raw.reshape((-1, n))
    lo = np.full(n, alpha)
    di = np.ones(n)
    up = np.full(n, alpha)
    lo[0] = up[0] = lo[-1] = up[-1] = 0.0
    cp = np.zeros(n)
    dp = np.zeros_like(r)
    cp[0] = up[0] / di[0]
    dp[:, 0] = r[:, 0] / di[0]
    for i in range(1, n):
        den = di[i] - lo[i] * cp[i - 1]
        cp[i] = up[i] / den if i < n - 1 else 0.0
        dp[:, i] = (r[:, i] - lo[i] * dp[:, i - 1]) / den
    out = np.zeros_like(dp)
    out[:, -1] = dp[:, -1]
    for i in range(n - 2, -1, -1):
        out[:, i] = dp[:, i] - cp[i] * out[:, i + 1]
    return out.reshape((*lead, n))


def fd(a, h, axis):
    """:268-275: central inside, one-sided at the two ends."""
    m = np.moveaxis(a, axis, -1)
    d = np.empty_like(m)
    d[..., 1:-1] = (m[..., 2:] - m[..., :-2]) / (2.0 * h)
    d[..., 0] = (m[..., 1] - m[..., 0]) / h
    d[..., -1] = (m[..., -1] - m[..., -2]) / h
    return np.moveaxis(d, -1, axis)


@dataclass(frozen=True)
class SensorParams:
    width: int = 4
    compression_threshold: float = 2.5
    jump_threshold: float = 0.04
    shear_threshold: float | None = 2.0
    boundary_guard: int = 4


def sensor(q, gas: Gas, dx, dy, sp: SensorParams):
    """:288-326: as the periodic sensor with one-sided ends; shear only for a
    positive threshold; a boundary band of width ``boundary_guard`` is set
    before the truncated dilation."""
    rho, u, v, p = primitives(q, gas)
    ux = fd(u, dx, -1)
    uy = fd(u, dy, -2)
    vx = fd(v, dx, -1)
    vy = fd(v, dy, -2)
    div = ux + vy
    div_rms = float(np.sqrt(np.mean(div**2)))
    comp = np.zeros_like(div, dtype=bool)
    if div_rms > 1e-12:
        comp = div < -sp.compression_threshold * div_rms
    shear = np.zeros_like(div, dtype=bool)
    if sp.shear_threshold is not None and sp.shear_threshold > 0.0:
        w = vx - uy
        w_rms = float(np.sqrt(np.mean(w**2)))
        if w_rms > 1e-12:
            shear = np.abs(w) > sp.shear_threshold * w_rms
    e = p / (rho * (gas.gamma - 1.0))
    jr = np.maximum(jump_indicator(rho, -1), jump_indicator(rho, -2))
    jp = np.maximum(jump_indicator(p, -1), jump_indicator(p, -2))
    je = np.maximum(jump_indicator(e, -1), jump_indicator(e, -2))
    raw = comp | shear | (jr > sp.jump_threshold) | (jp > sp.jump_threshold) | (je > sp.jump_threshold)
    g = sp.boundary_guard
    if g:
        raw[:g, :] = True
        raw[-g:, :] = True
        raw[:, :g] = True
        raw[:, -g:] = True
    return dilate(raw, sp.width)


def axis_derivative(q_ax, pf_ax, mask_ax, h, gas: Gas, normal: int, scheme: str):
    """:344-357."""
    rho, u, v, p = primitives(q_ax, gas)
    un = u if normal == 1 else v
    c = np.sqrt(gas.gamma * p / rho)
    alpha = float(np.max(np.abs(un) + c))
    smooth = smooth_flux(pf_ax)
    shock = compactify(weno_flux(q_ax, pf_ax, alpha))
    if scheme == "weno":
        hyb = shock
    else:
        hyb = np.where(iface_mask(mask_ax)[None, ...], shock, smooth)
    return compact_solve(hyb, h)


def rhs(q, gas: Gas, dx, dy, sp: SensorParams, scheme="hybrid"):
    """:359-369."""
    if scheme not in {"hybrid", "weno"}:
        raise ValueError("scheme must be 'hybrid' or 'weno'")
    qs = enforce(q, gas)
    m = sensor(qs, gas, dx, dy, sp)
    d_x = axis_derivative(qs, flux(qs, gas, "x"), m, dx, gas, 1, scheme)
    qt = np.moveaxis(qs, -2, -1)
    ft = np.moveaxis(flux(qs, gas, "y"), -2, -1)
    mt = np.moveaxis(m, -2, -1)
    d_y = np.moveaxis(axis_derivative(qt, ft, mt, dy, gas, 2, scheme), -1, -2)
    return -d_x - d_y


def fixed_dt(q, gas: Gas, dx, dy, cfl, t_end):
    """:371-378: dt from the state once; n = ceil(T/dt); dt = T/n."""
    rho, u, v, p = primitives(q, gas)
    c = np.sqrt(gas.gamma * p / rho)
    sr = (np.abs(u) + c) / dx + (np.abs(v) + c) / dy
    dt = cfl / float(np.max(sr))
    n = int(np.ceil(t_end / dt))
    return t_end / n, n


def _lap_edge(a, dx, dy):
    pad = np.pad(a, ((0, 0), (1, 1), (1, 1)), mode="edge")
    c = pad[:, 1:-1, 1:-1]
    lx = (pad[:, 1:-1, 2:] - 2.0 * c + pad[:, 1:-1, :-2]) / dx**2
    ly = (pad[:, 2:, 1:-1] - 2.0 * c + pad[:, :-2, 1:-1]) / dy**2
    return lx + ly


def hyperviscosity(q, compact_mask, gas: Gas, dx, dy, mn, weights=(1.0, 1.0, 1.0, 1.0)):
    """:381-416: edge-padded biharmonic filter on compact nodes, then enforce."""
    if compact_mask.shape != q.shape[-2:]:
        raise ValueError(f"compact_mask shape {compact_mask.shape} does not match grid {q.shape[-2:]}")
    h = min(dx, dy)
    bih = _lap_edge(_lap_edge(q, dx, dy), dx, dy)
    w = np.array([weights[0], weights[1], weights[2], weights[3]], dtype=q.dtype)[:, None, None]
    filtered = q - mn * w * h**4 * compact_mask[None, :, :] * bih
    return enforce(filtered, gas)


def outflow_guard(q, g=2):
    """:419-429: copy guard columns, then guard rows (corners follow rows)."""
    if g <= 0:
        return q
    out = np.array(q, copy=True)
    out[:, :, :g] = out[:, :, g][:, :, None]
    out[:, :, -g:] = out[:, :, -g - 1][:, :, None]
    out[:, :g, :] = out[:, g, :][:, None, :]
    out[:, -g:, :] = out[:, -g - 1, :][:, None, :]
    return out


def validate(q, gas: Gas, label="state"):
    """:432-452: raise FloatingPointError on NaN/Inf or rho, p <= 0 (raw q)."""
    if bool(np.any(~np.isfinite(q))):
        raise FloatingPointError(f"{label} contains NaN or Inf values")
    rho = q[0]
    u = q[1] / rho
    v = q[2] / rho
    ke = 0.5 * rho * (u**2 + v**2)
    p = (gas.gamma - 1.0) * (q[3] - ke)
    if bool(np.any(~np.isfinite(rho))) or bool(np.any(~np.isfinite(p))):
        raise FloatingPointError(f"{label} has non-finite density or pressure")
    if bool(np.any(rho <= 0.0)):
        raise FloatingPointError(f"{label} contains non-positive density")
    if bool(np.any(p <= 0.0)):
        raise FloatingPointError(f"{label} contains non-positive pressure")
