"""Forced 2-D HIT: initial condition, sources, Mach control and step loop.

Restates reference OOP/hit2d.py:139-204 (init), :264-276 (forcing source),
:279-349 (Mach control), :352-426 (drag, cooling), :429-548 (diagnostics),
:571-596 (snapshot fields), :599-615 (equation choice) and :764-918 (the
step loop, without file output).
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


def spectral_derivatives(a, nx, ny, dx, dy):
    """:146-155: d/dx, d/dy by fft2 -> i k -> ifft2().real."""
    kx, ky, _ = wavenumbers(nx, ny, dx, dy)
    ah = np.fft.fft2(a)
    return np.fft.ifft2(1j * kx * ah).real, np.fft.ifft2(1j * ky * ah).real


def divergence_and_vorticity(u, v, nx, ny, dx, dy):
    """:459-468."""
    du_dx, du_dy = spectral_derivatives(u, nx, ny, dx, dy)
    dv_dx, dv_dy = spectral_derivatives(v, nx, ny, dx, dy)
    return du_dx + dv_dy, dv_dx - du_dy


def large_scale_kinetic_fraction(q, gas, nx, ny, dx, dy, kmax):
    """:429-449."""
    if kmax <= 0.0:
        return 0.0
    rho, u, v, _ = primitives(q, gas)
    uf = u - np.mean(u)
    vf = v - np.mean(v)
    total = float(np.mean(rho * (uf**2 + vf**2)))
    if total <= np.finfo(float).eps:
        return 0.0
    ul = band_filter(uf, nx, ny, dx, dy, 0.0, kmax)
    vl = band_filter(vf, nx, ny, dx, dy, 0.0, kmax)
    return float(np.mean(rho * (ul**2 + vl**2))) / total


def source_power(s):
    """:452-456."""
    return float(np.mean(s[3]))


def compute_diagnostics(q, gas, nx, ny, dx, dy, length_x, length_y, initial_mass=None, forcing_info=None,
                        drag_power=0.0, cooling_power=0.0, large_scale_fraction=0.0, weno_fraction=0.0):
    """:471-548."""
    rho, u, v, pr = primitives(q, gas)
    div, vort = divergence_and_vorticity(u, v, nx, ny, dx, dy)
    c = np.sqrt(gas.gamma * pr / rho)
    mass = float(np.mean(rho) * length_x * length_y)
    uf = u - np.mean(u)
    vf = v - np.mean(v)
    uu = float(np.mean(rho * uf**2))
    vv = float(np.mean(rho * vf**2))
    uv = float(np.mean(rho * uf * vf))
    cs = uu + vv
    cov = np.sqrt(max(uu * vv, 0.0))
    d = {
        "kinetic_energy": 0.5 * cs,
        "Kx": 0.5 * uu,
        "Ky": 0.5 * vv,
        "A_K": abs(uu - vv) / cs if cs > 0.0 else 0.0,
        "C_uv": uv / cov if cov > 0.0 else 0.0,
        "rms_velocity": float(np.sqrt(np.mean(u**2 + v**2))),
        "turbulent_mach": float(np.sqrt(np.mean(u**2 + v**2)) / np.mean(c)),
        "mean_density": float(np.mean(rho)),
        "mean_pressure": float(np.mean(pr)),
        "divergence_rms": float(np.sqrt(np.mean(div**2))),
        "vorticity_rms": float(np.sqrt(np.mean(vort**2))),
        "mass": mass,
        "mass_error": 0.0 if initial_mass is None else mass - initial_mass,
        "large_scale_kinetic_fraction": large_scale_fraction,
        "drag_power": drag_power,
        "cooling_power": cooling_power,
        "weno_fraction": weno_fraction,
    }
    d.update({"P_in": 0.0, "Fxx": 0.0, "Fyy": 0.0, "Fxy": 0.0, "A_F": 0.0, "forcing_alpha": 0.0,
              "forcing_target_power": 0.0, "mach_control_mach": d["turbulent_mach"], "mach_control_target": 0.0,
              "mach_control_power_desired": 0.0})
    if forcing_info is not None:
        d.update({
            "P_in": forcing_info["injected_power"], "Fxx": forcing_info["Fxx"], "Fyy": forcing_info["Fyy"],
            "Fxy": forcing_info["Fxy"], "A_F": forcing_info["A_F"], "forcing_alpha": forcing_info["alpha"],
            "forcing_target_power": forcing_info.get("target_power", 0.0),
            "mach_control_mach": forcing_info.get("mach_control_mach", d["turbulent_mach"]),
            "mach_control_target": forcing_info.get("mach_control_target", 0.0),
            "mach_control_power_desired": forcing_info.get("mach_control_power_desired", 0.0),
        })
    return d


def snapshot_fields(q, gas, nx, ny, dx, dy):
    """:571-596: rho, u, v, pressure, vorticity, divergence."""
    rho, u, v, pr = primitives(q, gas)
    div, vort = divergence_and_vorticity(u, v, nx, ny, dx, dy)
    return {"rho": rho, "u": u, "v": v, "pressure": pr, "vorticity": vort, "divergence": div}


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

    def diagnostics(self, initial=False, initial_mass=None):
        """The loop's diagnostics of the current state: :806-818 (initial:
        no forcing information, no source powers) and :881-905."""
        p, gas, q = self.p, self.gas, self.q
        weno = float(np.mean(P.sensor(q, gas, p.dx, p.dy, p.sensor_params())))
        frac = large_scale_kinetic_fraction(q, gas, p.nx, p.ny, p.dx, p.dy, p.large_scale_drag_kmax)
        args = (q, gas, p.nx, p.ny, p.dx, p.dy, p.length, p.length)
        if initial:
            return compute_diagnostics(*args, large_scale_fraction=frac, weno_fraction=weno)
        drag = drag_source(q, gas, p.nx, p.ny, p.dx, p.dy, p.large_scale_drag, p.large_scale_drag_kmax)
        cool = cooling_source(q, gas, p.cooling_time, p.cooling_target_pressure)
        return compute_diagnostics(*args, initial_mass=initial_mass, forcing_info=self.info,
                                   drag_power=source_power(drag), cooling_power=source_power(cool),
                                   large_scale_fraction=frac, weno_fraction=weno)

    def run(self):
        while not self.done():
            self.step()
        return self.q, self.time, self.steps
