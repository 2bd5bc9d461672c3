"""Accuracy of the device's global sums in the FP32 mode.

The FP32 mode keeps the state, stencils and solves in single precision and
accumulates every global sum (time step and Mach control, sensor statistics,
forcing power and statistics) in double: a single-precision term converts to
double exactly, and a double sum of N terms is within about N * 2^-53 of the
exact sum of those terms (the per-thread runs and the block trees are much
shorter than N, so the error is far below that bound).  Against the FP32
error budget (1e-4 per field after 1000 steps, the state's own rounding is
6e-8 per step) that is negligible, which is why the sums are plain double
rather than compensated (Neumaier) double.

These tests make the sum's own error visible: the single-precision terms are
exact (products of values with at most 11 significant bits), spread over 2^44
in magnitude, so the only error left is the accumulation's.  The device sums
must agree with the exactly rounded sum (math.fsum) to 1e-12 of sum |t|
(one running double sum over all 512^2 terms would be off by 2e-13),
where a single-precision running sum of the same terms is off by more than
1e-9 of it.
"""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def few_bit_values(rng, shape, emin=-20, emax=2):
    """a * 2^e with a an integer in [1, 2047] and a random sign: squares and
    pairwise products are exact in float32."""
    a = rng.integers(1, 2048, size=shape).astype(np.float64)
    e = rng.integers(emin, emax + 1, size=shape)
    s = np.where(rng.random(shape) < 0.5, -1.0, 1.0)
    return (s * np.ldexp(a, e)).astype(np.float32)


def exact_sum(t):
    return math.fsum(np.asarray(t, np.float64).ravel().tolist())


def float32_running_sum(t):
    return float(np.add.accumulate(np.asarray(t, np.float32).ravel(), dtype=np.float32)[-1])


def plan32(n):
    from hit_solver_amd.runtime import PlanSpec, get_plan

    return get_plan(PlanSpec(nx=n, ny=n, dx=2 * np.pi / n, dy=2 * np.pi / n, periodic=True, dtype=torch.float32))


def check_sum(got, terms, what):
    """Error relative to sum |t| (the scale of an accumulation's rounding,
    whatever cancellation the signs bring)."""
    ref = exact_sum(terms)
    scale = exact_sum(np.abs(terms))
    err = abs(got - ref) / scale
    naive = abs(float32_running_sum(terms) - ref) / scale
    assert naive > 1e-9, f"{what}: the case does not separate single from double accumulation"
    assert err <= 1e-12, f"{what}: device sum rel err {err:.3e} (float32 running sum {naive:.3e})"


@pytest.mark.parametrize("n", [128, 512])
def test_step_reduce_sum_is_double_accurate(n):
    """sum(u^2 + v^2) of hitx_step_reduce (the Mach control's velocity sum)
    on an FP32 state with rho = 1 (u = q1 exactly) and v = 0."""
    rng = np.random.default_rng(n)
    u = few_bit_values(rng, (n, n))
    q = np.zeros((4, n, n), np.float32)
    q[0] = 1.0
    q[1] = u
    q[3] = 1.0e3  # p = 0.4 (E - u^2/2) > 0: no floor applies
    p = plan32(n)
    _, s_uv, _ = p.step_reduce(torch.from_numpy(q).cuda())
    check_sum(s_uv, u.astype(np.float64) ** 2, f"step_reduce {n}^2")


@pytest.mark.parametrize("n", [128, 512])
def test_forcing_statistics_are_double_accurate(n):
    """sum(fx^2), sum(fy^2), sum(fx fy) of hitx_forcing_scale with scale =
    alpha = 1 (the forcing statistics of one step)."""
    rng = np.random.default_rng(7 * n)
    fx = few_bit_values(rng, (n, n))
    fy = few_bit_values(rng, (n, n), -16, 0)
    p = plan32(n)
    d = {k: torch.from_numpy(np.ascontiguousarray(a)).cuda() for k, a in
         (("fx", fx), ("fy", fy), ("rho", np.ones((n, n), np.float32)), ("u", np.zeros((n, n), np.float32)),
          ("v", np.zeros((n, n), np.float32)))}
    st = p.forcing_scale(d["fx"], d["fy"], 1.0, 1.0, d["rho"], d["u"], d["v"]).cpu().numpy()
    np.testing.assert_array_equal(d["fx"].cpu().numpy(), fx)  # scale 1, alpha 1: unchanged
    X, Y = fx.astype(np.float64), fy.astype(np.float64)
    check_sum(float(st[0]), X * X, "sum fx^2")
    check_sum(float(st[1]), Y * Y, "sum fy^2")
    check_sum(float(st[2]), X * Y, "sum fx fy")
    assert float(st[3]) == 0.0
