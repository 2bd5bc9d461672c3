"""Solenoidal OU shell forcing on the hip backend (reference OOP/forcing.py:10-130).

The noise stream is NumPy PCG64 ``default_rng(seed).standard_normal``, as on
the reference's NumPy path (the reference's CuPy path draws a different
stream, OOP/forcing.py:67-68), generated bit for bit on the device
(``hitx_normals``).  Everything grid-sized runs on the device: the noise, the
forward DFT of the noise restricted to the shell modes
(``hitx_forcing_project``), the curl and inverse DFT (``hitx_forcing_synthesize``),
the rescale and the statistics (``hitx_forcing_scale``).  The OU recurrence
and the power-rescale logic act on the ~56 shell coefficients and scalars on
the host, in the reference's operation order.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .domain import Domain2D
from .rng import DeviceNormalStream
from .runtime import PlanSpec, get_plan


def shell_modes(domain: Domain2D, k_min: float, k_max: float):
    """Index pairs (kx index, ky index) of the shell, with the physical
    wavenumbers of OOP/forcing.py:47-59 (same float expressions)."""
    kxv = 2.0 * np.pi * np.fft.fftfreq(domain.nx, d=domain.dx)
    kyv = 2.0 * np.pi * np.fft.fftfreq(domain.ny, d=domain.dy)
    kx, ky = np.meshgrid(kxv, kyv, indexing="xy")
    mag = np.sqrt(kx**2 + ky**2)
    mask = (mag >= float(k_min)) & (mag <= float(k_max)) & (mag > 0.0)
    iy, ix = np.nonzero(mask)
    fx_idx = np.fft.fftfreq(domain.nx, d=1.0 / domain.nx).round().astype(int)
    fy_idx = np.fft.fftfreq(domain.ny, d=1.0 / domain.ny).round().astype(int)
    idx = np.stack([fx_idx[ix], fy_idx[iy]], axis=1)
    return idx, kx[iy, ix], ky[iy, ix]


@dataclass
class IsotropicShellOUForcing2D:
    domain: Domain2D
    k_min: float
    k_max: float
    correlation_time: float = 1.0
    force_rms: float = 1.0
    target_power: float | None = 1.0e-3
    min_power: float = 1.0e-6
    max_rescale: float | None = 20.0
    alpha_memory: float = 0.2
    seed: int = 1234
    xp: object = None
    dtype: torch.dtype = torch.float64
    _phi: object = field(default=None, init=False, repr=False)
    _alpha: float = field(default=1.0, init=False, repr=False)
    _rng: object = field(default=None, init=False, repr=False)

    def __post_init__(self):
        if self.k_min < 0.0 or self.k_max <= 0.0 or self.k_min > self.k_max:
            raise ValueError("forcing shell must satisfy 0 <= k_min <= k_max")
        if self.correlation_time <= 0.0:
            raise ValueError("forcing correlation time must be positive")
        if not 0.0 <= self.alpha_memory < 1.0:
            raise ValueError("alpha_memory must be in [0, 1)")
        self.modes, self._kx, self._ky = shell_modes(self.domain, self.k_min, self.k_max)
        if len(self.modes) == 0:
            raise ValueError("forcing shell contains no Fourier modes on this grid")
        if np.abs(self.modes).max() > 12 or len(self.modes) > 116:
            raise ValueError("forcing shell too wide for the device projection (|k index| <= 12)")
        self._phi = np.zeros(len(self.modes), dtype=np.complex128)
        self._rng = DeviceNormalStream(self.seed)
        self._noise = None

    @property
    def mode_tuple(self):
        return tuple(int(v) for v in self.modes.reshape(-1))

    def plan(self):
        d = self.domain
        return get_plan(PlanSpec(nx=d.nx, ny=d.ny, dx=d.dx, dy=d.dy, periodic=True, dtype=self.dtype,
                                 force_modes=self.mode_tuple))

    def draw_noise(self):
        """The next nx*ny normals of the PCG64 stream (float64, on the device)."""
        if self._noise is None:
            self._noise = torch.empty((self.domain.ny, self.domain.nx), dtype=torch.float64, device="cuda")
        return self._rng.normals(self._noise.shape, out=self._noise)

    def update(self, dt: float, rho, u, v):
        """OOP/forcing.py:70-130: one OU step; returns (fx, fy, info)."""
        if dt <= 0.0:
            raise ValueError("forcing update requires dt > 0")
        plan = self.plan()
        decay = float(np.exp(-dt / self.correlation_time))
        incr = float(np.sqrt(max(1.0 - decay**2, 0.0)))
        noise = self.draw_noise()
        r = plan.forcing_project(noise).cpu().numpy()
        rhat = r[0::2] + 1j * r[1::2]
        self._phi = decay * self._phi + incr * rhat
        fxh = 1j * self._ky * self._phi
        fyh = -1j * self._kx * self._phi
        fhat = np.concatenate([np.stack([fxh.real, fxh.imag], 1).ravel(),
                               np.stack([fyh.real, fyh.imag], 1).ravel()])
        fhat_d = torch.from_numpy(fhat).to("cuda")
        shp = (self.domain.ny, self.domain.nx)
        fx = torch.empty(shp, dtype=self.dtype, device="cuda")
        fy = torch.empty(shp, dtype=self.dtype, device="cuda")
        s = plan.forcing_synthesize(fhat_d, fx, fy, rho, u, v).cpu().numpy()
        N = float(self.domain.nx * self.domain.ny)
        vrms = float(np.sqrt(s[0] / N))
        if self.force_rms <= 0.0:
            scale = 0.0
        elif vrms > np.finfo(float).eps:
            scale = self.force_rms / vrms
        else:
            scale = 1.0
        p_before = float(scale * s[1] / N)
        a_tgt = self._alpha
        if self.target_power is not None and abs(p_before) > self.min_power:
            a_tgt = self.target_power / p_before
            if self.max_rescale is not None:
                a_tgt = float(np.clip(a_tgt, -self.max_rescale, self.max_rescale))
            self._alpha = self.alpha_memory * self._alpha + (1.0 - self.alpha_memory) * a_tgt
        elif self.target_power is None:
            self._alpha = 1.0
        st = plan.forcing_scale(fx, fy, scale, self._alpha, rho, u, v).cpu().numpy()
        fxx, fyy, fxy = float(st[0] / N), float(st[1] / N), float(st[2] / N)
        den = fxx + fyy
        return fx, fy, {
            "alpha": float(self._alpha),
            "alpha_target": float(a_tgt),
            "target_power": float("nan") if self.target_power is None else float(self.target_power),
            "power_before_rescale": p_before,
            "injected_power": float(st[3] / N),
            "Fxx": fxx,
            "Fyy": fyy,
            "Fxy": fxy,
            "A_F": abs(fxx - fyy) / den if den > 0.0 else 0.0,
            "ou_decay": decay,
        }
