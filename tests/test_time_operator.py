"""TimeOperator and the generic SSPRK3 branch (reference OOP/time_operator.py:17-69).

CPU: the loop's bookkeeping against closed forms on dq/dt = lam q, whose
SSP-RK3 step multiplies by 1 + z + z^2/2 + z^3/6 (z = lam dt): the step
count ceil(T/dt) with dt reset to T/n, the saved times and states, the
state filter applied after every stage and once at construction, and hooks
applied after the step and cleaned again.
GPU: the loop driving the device operator RHS (generic callable path,
enforce as the state filter) against the same loop over the oracle's RHS.
"""

import math

import numpy as np
import pytest
import torch

from hit_solver_amd.time_operator import SSPRK3, TimeOperator


def test_time_operator_linear_closed_form():
    lam = -0.7
    q0 = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    op = TimeOperator(state=q0.clone(), rhs=lambda q: lam * q, dt=0.3)
    times, states = op.run(1.0, save_every=2)
    n = math.ceil(1.0 / 0.3)
    assert n == 4 and op.dt == pytest.approx(0.25, rel=0, abs=0)
    np.testing.assert_allclose(times.numpy(), [0.0, 0.5, 1.0])
    z = lam * 0.25
    g = 1 + z + z * z / 2 + z**3 / 6
    for k, s in zip((0, 2, 4), states):
        np.testing.assert_allclose(s.numpy(), q0.numpy() * g**k, rtol=1e-14)


def test_time_operator_filter_and_hooks():
    calls = []

    def filt(q):
        calls.append("f")
        return torch.clamp(q, min=0.0)

    def hook(i, q):
        calls.append(f"h{i}")
        return q + 1.0

    op = TimeOperator(state=torch.tensor([-1.0, 2.0], dtype=torch.float64), rhs=lambda q: -q, dt=0.1,
                      state_filter=filt, step_hooks=[hook])
    assert op.state[0].item() == 0.0  # filtered at construction
    op.step(0)
    # three stage cleans, the hook, then one more clean
    assert calls == ["f", "f", "f", "f", "h0", "f"]
    g = 1 - 0.1 + 0.01 / 2 - 0.001 / 6
    np.testing.assert_allclose(op.state.numpy(), [1.0, 2.0 * g + 1.0], rtol=1e-14)


def test_ssprk3_generic_branch_matches_shu_osher():
    q0 = torch.tensor([0.3, -1.2], dtype=torch.float64)
    f = lambda q: torch.sin(q)  # noqa: E731
    dt = 0.05
    q1 = q0 + dt * f(q0)
    q2 = 0.75 * q0 + 0.25 * (q1 + dt * f(q1))
    ref = (1.0 / 3.0) * q0 + (2.0 / 3.0) * (q2 + dt * f(q2))
    np.testing.assert_array_equal(SSPRK3().step(q0, f, dt).numpy(), ref.numpy())


@pytest.mark.gpu
def test_time_operator_device_rhs_vs_oracle():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from hit_solver_amd.domain import Domain2D
    from hit_solver_amd.equations import EulerEquation2D
    from hit_solver_amd.spatial_operator import PeriodicEulerShockSensor2D, PeriodicHybridEuler2DOperator
    from oracle import hit as OH
    from oracle import periodic as OP
    from oracle.equations import Gas, enforce
    from oracle.rk import ssprk3

    nx, ny = 64, 48
    p = OH.HITParams(nx=nx, ny=ny, target_mach=0.3, initial_kmin=3, initial_kmax=5)
    q = OH.initial_state(p, np.random.default_rng(17))
    d = Domain2D(0.0, 2 * np.pi, 0.0, 2 * np.pi, nx, ny)
    eq = EulerEquation2D(gamma=1.4)
    op = PeriodicHybridEuler2DOperator(d, eq, PeriodicEulerShockSensor2D(d, eq, width=4, jump_threshold=0.04))
    dt = 0.9 * op.stable_time_step(torch.from_numpy(q).cuda(), 0.05, 1.0)
    t_end = 5.5 * dt
    loop = TimeOperator(state=torch.from_numpy(q).cuda(), rhs=op.rhs, dt=dt, state_filter=eq.enforce_physical_state)
    times, states = loop.run(t_end, save_every=3)
    assert len(states) == 3 and times[-1].item() == pytest.approx(t_end, rel=1e-15)
    gas, sp = Gas(), OP.SensorParams()
    r = enforce(q, gas)
    n = math.ceil(t_end / dt)
    for _ in range(n):
        r = ssprk3(r, lambda s: OP.rhs(s, gas, p.dx, p.dy, sp), t_end / n, clean=lambda s: enforce(s, gas))
    got = states[-1].cpu().numpy()
    for k in range(4):
        e = float(np.linalg.norm(got[k] - r[k]) / np.linalg.norm(r[k]))
        assert e <= 1e-12, (k, e)
