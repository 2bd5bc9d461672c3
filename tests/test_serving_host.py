"""Host logic of the serving loop (no GPU): request configurations, the
Poisson stream, kernel classes."""

import math

import numpy as np

from oracle import riemann as OR


def test_riemann_request_length_matches_fixed_dt():
    """tfinal = steps * dt0 with dt0 from the quadrant states gives the
    oracle's fixed-dt loop (OOP/problems/riemann_config3.py:371-378) that many
    steps."""
    from hit_solver_amd.serving import Request, request_config

    for kind, fn in (("c3", OR.config3), ("c3_08", OR.config3_08), ("c6", OR.config6)):
        for n, steps in ((64, 200), (96, 1234)):
            cfg = request_config(Request(0, kind, n, "fp64", steps, 0.0))
            p = fn(nx=n, ny=n, tfinal=cfg.tfinal)
            run = OR.RiemannRun(p)
            assert run.n_steps in (steps, steps + 1), (kind, n, run.n_steps)


def test_poisson_stream_statistics():
    from hit_solver_amd.serving import KINDS, poisson_stream

    reqs = poisson_stream(4000, rate=20.0, seed=3)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 20.0) < 0.1 / 20.0
    assert {r.kind for r in reqs} == set(KINDS)
    assert min(r.n for r in reqs) == 128 and max(r.n for r in reqs) == 1024
    assert all(200 <= r.steps <= 5000 for r in reqs)
    assert {r.dtype for r in reqs} == {"fp32", "fp64"}


def test_class_keys_group_kernel_compatible_requests():
    """One class per element type: HIT and the Riemann problems share a batch
    (the batch runs each instance as in a batch of its own kind)."""
    from hit_solver_amd.serving import Request, class_key, request_config

    k = lambda kind, dt: class_key(request_config(Request(0, kind, 128, dt, 100, 0.0)))  # noqa: E731
    assert k("c3", "fp64") == k("c6", "fp64") == k("c3_08", "fp64") == k("hit", "fp64")
    assert k("c3", "fp64") != k("c3", "fp32")
    assert k("hit", "fp32") == k("c3", "fp32")
    assert math.isfinite(request_config(Request(0, "c6", 256, "fp32", 300, 0.0)).tfinal)
