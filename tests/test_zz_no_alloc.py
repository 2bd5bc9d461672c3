"""hitx_rk3_stage_static: the stage entry point that allocates nothing (its
workspace is a static device array) against hitx_rk3_stage with a caller
workspace.  Same kernels on the same inputs: bit-for-bit equal outputs, on
consecutive calls with a smaller grid second (the shared array keeps the
first call's contents), for the forced-HIT closure and the Riemann guard
path, in FP64 and FP32; a grid whose workspace exceeds the array is rejected
before any device work."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _hit_case(n, dtype, seed):
    from hit_solver_amd.domain import Domain2D
    from hit_solver_amd.equations import CompressibleNavierStokes2D
    from hit_solver_amd.hit2d import drag_modes
    from hit_solver_amd.spatial_operator import PeriodicEulerShockSensor2D, PeriodicHybridEuler2DOperator
    from oracle import hit as OH

    p = OH.HITParams(nx=n, ny=n, target_mach=0.3, initial_kmin=3, initial_kmax=5)
    q = OH.initial_state(p, np.random.default_rng(seed))
    x = np.arange(n) * p.dx
    q[0] *= 1.0 + 0.5 * (np.sin(x)[None, :] > 0.4)  # density steps: WENO bands
    d = Domain2D(0.0, 2 * np.pi, 0.0, 2 * np.pi, n, n)
    eq = CompressibleNavierStokes2D(gamma=1.4, viscosity=7.5e-4)
    op = PeriodicHybridEuler2DOperator(d, eq, PeriodicEulerShockSensor2D(d, eq, width=4, jump_threshold=0.04))
    dm = drag_modes(d, 2.0)
    plan = op.plan(dtype, viscosity=7.5e-4, forcing=True, drag_coef=0.10,
                   drag_modes=tuple(int(v) for v in dm.reshape(-1)), cooling_time=5.0)
    g = np.random.default_rng(seed + 1)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)  # noqa: E731
    return plan, dev(q), dev(0.1 * g.standard_normal((n, n))), dev(0.1 * g.standard_normal((n, n)))


def _riemann_case(n, dtype):
    from hit_solver_amd.riemann import RiemannConfig3, RiemannSolver

    s = RiemannSolver(RiemannConfig3(nx=n, ny=n, shear_threshold=None, dtype="fp64" if dtype == torch.float64
                                     else "fp32", progress_every=10**9))
    s.run(max_steps=3)
    return s.rhs.plan, s.q.clone(), None, None


def _stage_both(plan, q, fx, fy, dt):
    """Three chained stages through each entry point; returns both triples."""
    from hit_solver_amd._lib import check, lib

    p = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    outs = {}
    for which in ("ws", "static"):
        bufs = [torch.full_like(q, float("nan")) for _ in range(3)]
        src = q
        for st in range(3):
            if which == "ws":
                check(lib().hitx_rk3_stage(plan.handle, st, p(q), p(src), p(bufs[st]), p(fx), p(fy), float(dt),
                                           p(plan.ws), plan.ws_bytes, stream))
            else:
                check(lib().hitx_rk3_stage_static(plan.handle, st, p(q), p(src), p(bufs[st]), p(fx), p(fy),
                                                  float(dt), stream))
            src = bufs[st]
        torch.cuda.synchronize()
        outs[which] = [b.cpu().numpy() for b in bufs]
    return outs


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_static_stage_equals_workspace_stage(dtype):
    cases = [("HIT 96^2", lambda: _hit_case(96, dtype, 3)), ("HIT 64^2", lambda: _hit_case(64, dtype, 4)),
             ("Riemann 80^2", lambda: _riemann_case(80, dtype)), ("Riemann 48^2", lambda: _riemann_case(48, dtype))]
    for name, make in cases:
        print(f"[no-alloc] {name} {dtype}", flush=True)
        plan, q, fx, fy = make()
        outs = _stage_both(plan, q, fx, fy, 1.0e-3)
        for st in range(3):
            assert np.isfinite(outs["ws"][st]).all()
            np.testing.assert_array_equal(outs["static"][st], outs["ws"][st], err_msg=f"{name} stage {st}")


def test_static_stage_rejects_large_grid():
    from hit_solver_amd._lib import check, lib

    plan, q, fx, fy = _hit_case(1024, torch.float64, 5)
    assert plan.ws_bytes > 32 << 20
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    out = torch.empty_like(q)
    with pytest.raises(ValueError, match="static scratch"):
        check(lib().hitx_rk3_stage_static(plan.handle, 0, p(q), p(q), p(out), p(fx), p(fy), 1.0e-3,
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
