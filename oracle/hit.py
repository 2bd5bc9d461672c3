"""Forced 2-D HIT: initial condition, sources, Mach control and step loop.

Restates reference OOP/hit2d.py:139-204 (init), :264-276 (forcing source),
:279-349 (Mach control), :352-426 (drag, cooling), :599-615 (equation
choice) and :764-918 (the step loop, without file output).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import periodic as P
from .equations import Gas, conservative, enforce, primitives, viscous_rhs
from .forcing import ShellOUForcing
from .rk import ssprk3


@dataclass(frozen=True)
class HITParams:
    """Mirror of the reference HIT2DConfig fields used on the hot path
    (OOP/hit2d.py:36-80)."""

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
    max_steps: int | None = None

    @property
    def dx(self):
        return (self.length - 0.0) / self.nx

    @property
    def dy(self):
        return (self.length - 0.0) / self.ny

    def gas(self):
        return Gas(gamma=self.gamma, viscosity=self.viscosity, prandtl=self.prandtl)

    def sensor_params(self):
        return P.SensorParams(self.sensor_width, self.compression_threshold,
                              self.jump_threshold, self.shear_threshold)


def wavenumbers(nx, ny, dx, dy):
    """:139-143."""
    kx, ky = np.meshgrid(2.0 * np.pi * np.fft.fftfreq(nx, d=dx),
                         2.0 * np.pi * np.fft.fftfreq(ny, d=dy), indexing="xy")
    return kx, ky, kx**2 + ky**2


def mesh(p: HITParams):
    """OOP/domain.py:69-87: cell centres, arrays [ny, nx]."""
    x = 0.0 + p.dx * (np.arange(p.nx) + 0.5)
    y = 0.0 + p.dy * (np.arange(p.ny) + 0.5)
    return np.meshgrid(x, y, indexing="xy")


def initial_state(p: HITParams, rng):
    """:158-204: random-phase streamfunction on the half-plane of the shell,
    amplitude 1/k^2, spectral curl, mean removal, rho = 1, p = 1/gamma,
    velocity scaled to the target turbulent Mach number."""
    X, Y = mesh(p)
    psi = np.zeros_like(X)
    for kx in range(-p.initial_kmax, p.initial_kmax + 1):
        for ky in range(-p.initial_kmax, p.initial_kmax + 1):
            if ky < 0 or (ky == 0 and kx <= 0):
                continue
            mag = np.sqrt(kx**2 + ky**2)
            if mag < p.initial_kmin or mag > p.initial_kmax:
                continue
            amp = 1.0 / max(kx**2 + ky**2, 1)
            psi += amp * np.cos(kx * X + ky * Y + rng.uniform(0.0, 2.0 * np.pi))
    KX, KY, _ = wavenumbers(p.nx, p.ny, p.dx, p.dy)
    ph = np.fft.fft2(psi)
    u = np.fft.ifft2(1j * KY * ph).real
    v = -np.fft.ifft2(1j * KX * ph).real
    u -= np.mean(u)
    v -= np.mean(v)
    rho = np.ones_like(u)
    pr = np.full_like(u, 1.0 / p.gamma)
    c = np.sqrt(p.gamma * pr / rho)
    vrms = np.sqrt(np.mean(u**2 + v**2))
    if float(vrms) <= np.finfo(float).eps:
        raise ValueError("initial streamfunction produced zero velocity")
    scale = p.target_mach * float(np.mean(c)) / vrms
    u *= scale
    v *= scale
    return conservative(rho, u, v, pr, Gas(gamma=p.gamma))


def forcing_source(q, gas, fx, fy):
    """:264-276."""
    rho, u, v, _ = primitives(q, gas)
    s = np.zeros_like(q)
    s[1] = rho * fx
    s[2] = rho * fy
    s[3] = rho * (u * fx + v * fy)
    return s


def band_filter(a, nx, ny, dx, dy, k_min=0.0, k_max=None):
    """:352-367."""
    kx, ky, k2 = wavenumbers(nx, ny, dx, dy)
    mag = np.sqrt(k2)
    m = (mag >= float(k_min)) & (k2 > 0.0)
    if k_max is not None:
        m = m & (mag <= float(k_max))
    return np.fft.ifft2(np.where(m, np.fft.fft2(a), 0.0)).real


def drag_source(q, gas, nx, ny, dx, dy, coef, kmax):
    """:370-399: -coef * (0 < |k| <= kmax part of the momentum fluctuation)."""
    s = np.zeros_like(q)
    if coef <= 0.0 or kmax <= 0.0:
        return s
    _, u, v, _ = primitives(q, gas)
    mx = q[1] - np.mean(q[1])
    my = q[2] - np.mean(q[2])
    lx = band_filter(mx, nx, ny, dx, dy, 0.0, kmax)
    ly = band_filter(my, nx, ny, dx, dy, 0.0, kmax)
    s[1] = -coef * lx
    s[2] = -coef * ly
    s[3] = u * s[1] + v * s[2]
    return s


def cooling_source(q, gas, tc, p_ref=None):
    """:402-426: uniform -(<p> - p_ref)/((gamma-1) t_c) energy sink."""
    s = np.zeros_like(q)
    if tc is None or tc <= 0.0:
        return s
    _, _, _, pr = primitives(q, gas)
    ref = 1.0 / gas.gamma if p_ref is None else p_ref
    err = (np.mean(pr) - ref) / (gas.gamma - 1.0)
    s[3] = -err / tc
    return s


def turbulent_mach(rho, u, v, pr, gamma):
    """:279-290."""
    c = np.sqrt(gamma * pr / rho)
    return float(np.sqrt(np.mean(u**2 + v**2)) / np.mean(c))


def mach_control(p: HITParams, power, rho, u, v, pr):
    """:293-349: slow feedback on the target power towards a target Mt."""
    tgt = p.target_mach if p.mach_control_target is None else p.mach_control_target
    mt = turbulent_mach(rho, u, v, pr, p.gamma)
    desired = np.nan if power is None else float(power)
    if not p.mach_control or power is None:
        return power, {"mach_control_mach": mt, "mach_control_target": float(tgt),
                       "mach_control_power_desired": desired}
    if tgt <= 0.0:
        raise ValueError("Mach control requires a positive target Mach number")
    if not 0.0 <= p.mach_control_memory < 1.0:
        raise ValueError("mach_control_memory must be in [0, 1)")
    if p.mach_control_exponent <= 0.0:
        raise ValueError("mach_control_exponent must be positive")
    ratio = (tgt / max(mt, 1.0e-8)) ** p.mach_control_exponent
    desired = float(power * ratio)
    if p.mach_control_min_power is not None:
        desired = max(desired, p.mach_control_min_power)
    if p.mach_control_max_power is not None:
        desired = min(desired, p.mach_control_max_power)
    nxt = p.mach_control_memory * power + (1.0 - p.mach_control_memory) * desired
    return nxt, {"mach_control_mach": mt, "mach_control_target": float(tgt),
                 "mach_control_power_desired": desired}


def make_forcing(p: HITParams):
    """:791-803 (seed + 1)."""
    return ShellOUForcing(
        p.nx, p.ny, p.dx, p.dy,
        k_min=0.0 if p.forcing_kmin is None else p.forcing_kmin, k_max=p.forcing_kmax,
        correlation_time=p.forcing_correlation_time, force_rms=p.forcing_rms,
        target_power=p.target_energy_injection, min_power=p.min_forcing_power,
        max_rescale=p.max_forcing_rescale, alpha_memory=p.forcing_alpha_memory,
        seed=p.forcing_seed + 1,
    )


def rhs_closure(p: HITParams, gas, fx, fy):
    """:847-867: operator rhs + viscous + forcing + drag + cooling."""
    sp = p.sensor_params()

    def L(state):
        r = P.rhs(state, gas, p.dx, p.dy, sp)
        if p.viscosity > 0.0:
            r = r + viscous_rhs(state, gas, p.dx, p.dy)
        s = r + forcing_source(state, gas, fx, fy)
        if p.large_scale_drag > 0.0:
            s = s + drag_source(state, gas, p.nx, p.ny, p.dx, p.dy, p.large_scale_drag,
                                p.large_scale_drag_kmax)
        if p.cooling_time is not None and p.cooling_time > 0.0:
            s = s + cooling_source(state, gas, p.cooling_time, p.cooling_target_pressure)
        return s

    return L


class HITRun:
    """The reference step loop (:764-918) as a resumable object, without
    diagnostics or file output.  ``step()`` advances one RK3 step."""

    def __init__(self, p: HITParams, q0=None):
        if p.hyperviscosity_interval <= 0 and p.hyperviscosity_mn > 0.0:
            raise ValueError("hyperviscosity_interval must be positive when mn > 0")
        self.p = p
        self.gas = p.gas()
        rng = np.random.default_rng(p.forcing_seed)
        self.forcing = make_forcing(p)
        q = initial_state(p, rng) if q0 is None else q0
        self.q = enforce(q, self.gas)
        self.time = 0.0
        self.steps = 0
        self.power = p.target_energy_injection
        self.info = None

    def done(self):
        if self.p.max_steps is not None and self.steps >= self.p.max_steps:
            return True
        return not self.time < self.p.tfinal

    def step(self):
        p, gas = self.p, self.gas
        q = self.q
        dt = P.stable_dt(q, gas, p.dx, p.dy, p.cfl, p.tfinal - self.time)
        if p.viscosity > 0.0:
            h = min(p.dx, p.dy)
            dt = min(dt, 0.25 * h**2 / p.viscosity, p.tfinal - self.time)
        rho, u, v, pr = primitives(q, gas)
        self.power, minfo = mach_control(p, self.power, rho, u, v, pr)
        self.forcing.target_power = self.power
        fx, fy, info = self.forcing.update(dt, rho, u, v)
        info.update(minfo)
        self.info = info
        q = ssprk3(q, rhs_closure(p, gas, fx, fy), dt, clean=lambda s: enforce(s, gas))
        self.steps += 1
        self.time += dt
        if p.hyperviscosity_mn > 0.0 and self.steps % p.hyperviscosity_interval == 0:
            m = P.sensor(q, gas, p.dx, p.dy, p.sensor_params())
            q = P.hyperviscosity(q, gas, p.dx, p.dy, p.hyperviscosity_mn, ~m)
        self.q = q
        self.last_dt = dt
        return q

    def run(self):
        while not self.done():
            self.step()
        return self.q, self.time, self.steps
