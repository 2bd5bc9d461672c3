"""Solenoidal Ornstein-Uhlenbeck shell forcing, reference OOP/forcing.py:10-130.

Noise is NumPy PCG64 ``default_rng(seed).standard_normal((ny, nx))`` (the
reference's NumPy branch, :65-66, :77); the CuPy branch's different stream
is not reproduced.
"""

from __future__ import annotations

import numpy as np


class ShellOUForcing:
    def __init__(self, nx, ny, dx, dy, k_min, k_max, correlation_time=1.0, force_rms=1.0,
                 target_power=1.0e-3, min_power=1.0e-6, max_rescale=20.0, alpha_memory=0.2,
                 seed=1234):
        # :39-68
        if k_min < 0.0 or k_max <= 0.0 or k_min > k_max:
            raise ValueError("forcing shell must satisfy 0 <= k_min <= k_max")
        if correlation_time <= 0.0:
            raise ValueError("forcing correlation time must be positive")
        if not 0.0 <= alpha_memory < 1.0:
            raise ValueError("alpha_memory must be in [0, 1)")
        self.nx, self.ny = nx, ny
        self.correlation_time = correlation_time
        self.force_rms = force_rms
        self.target_power = target_power
        self.min_power = min_power
        self.max_rescale = max_rescale
        self.alpha_memory = alpha_memory
        kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
        ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
        self.kx, self.ky = np.meshgrid(kx, ky, indexing="xy")
        mag = np.sqrt(self.kx**2 + self.ky**2)
        self.mask = (mag >= float(k_min)) & (mag <= float(k_max)) & (mag > 0.0)
        if int(np.count_nonzero(self.mask)) == 0:
            raise ValueError("forcing shell contains no Fourier modes on this grid")
        self.phi_hat = np.zeros((ny, nx), dtype=np.complex128)
        self.alpha = 1.0
        self.rng = np.random.default_rng(seed)

    def update(self, dt, rho, u, v):
        """:70-130: one OU step; returns (fx, fy, info)."""
        if dt <= 0.0:
            raise ValueError("forcing update requires dt > 0")
        decay = float(np.exp(-dt / self.correlation_time))
        incr = float(np.sqrt(max(1.0 - decay**2, 0.0)))
        noise = self.rng.standard_normal((self.ny, self.nx))
        nh = np.where(self.mask, np.fft.fft2(noise), 0.0)
        self.phi_hat = decay * self.phi_hat + incr * nh
        fx = np.fft.ifft2(1j * self.ky * self.phi_hat).real
        fy = np.fft.ifft2(-1j * self.kx * self.phi_hat).real
        vrms = float(np.sqrt(np.mean(fx**2 + fy**2)))
        if self.force_rms <= 0.0:
            fx.fill(0.0)
            fy.fill(0.0)
        elif vrms > np.finfo(float).eps:
            s = self.force_rms / vrms
            fx *= s
            fy *= s
        p_before = float(np.mean(rho * (fx * u + fy * v)))
        a_tgt = self.alpha
        if self.target_power is not None and abs(p_before) > self.min_power:
            a_tgt = self.target_power / p_before
            if self.max_rescale is not None:
                a_tgt = float(np.clip(a_tgt, -self.max_rescale, self.max_rescale))
            self.alpha = self.alpha_memory * self.alpha + (1.0 - self.alpha_memory) * a_tgt
        elif self.target_power is None:
            self.alpha = 1.0
        fx *= self.alpha
        fy *= self.alpha
        fxx = float(np.mean(fx**2))
        fyy = float(np.mean(fy**2))
        fxy = float(np.mean(fx * fy))
        den = fxx + fyy
        a_f = abs(fxx - fyy) / den if den > 0.0 else 0.0
        p_in = float(np.mean(rho * (fx * u + fy * v)))
        return fx, fy, {
            "alpha": float(self.alpha),
            "alpha_target": float(a_tgt),
            "target_power": float("nan") if self.target_power is None else float(self.target_power),
            "power_before_rescale": p_before,
            "injected_power": p_in,
            "Fxx": fxx,
            "Fyy": fyy,
            "Fxy": fxy,
            "A_F": a_f,
            "ou_decay": decay,
        }
