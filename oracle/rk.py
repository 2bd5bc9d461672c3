"""Shu-Osher SSP-RK3, reference OOP/time_operator.py:17-26."""

from __future__ import annotations


def ssprk3(q0, L, dt, clean=None):
    c = clean if clean is not None else (lambda s: s)
    q1 = c(q0 + dt * L(q0))
    q2 = c(0.75 * q0 + 0.25 * (q1 + dt * L(q1)))
    return c((1.0 / 3.0) * q0 + (2.0 / 3.0) * (q2 + dt * L(q2)))
