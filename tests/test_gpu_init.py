"""Initial states built on the device vs the reference's golden vectors and
the oracle.

* HIT (OOP/hit2d.py:158-204): the reference differentiates the
  streamfunction by FFT, the device by the equivalent periodic convolution,
  and the device's cos may differ from the host libm's in the last bit, so
  the state agrees to rounding: per-field relative L2 <= 1e-13 against the
  reference's own 32x24 state (golden F6), and the reference's diagnostics
  of its 512^2 initial state (kinetic energy, Mt, vorticity rms to 1e-12;
  the divergence, rounding noise, to 1e-12 absolute).  The WENO mask of a
  state that differs by rounding may differ at nodes whose jump indicator
  sits at the threshold: the count is compared within one dilated node.
* Riemann quadrant states: bit-exact (the same arithmetic).
* FP32 states are the FP64 states rounded once.
"""

import numpy as np
import pytest
import torch

from oracle import riemann as OR

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _hit(nx, ny, dtype=torch.float64, seed=1234):
    from hit_solver_amd.equations import EulerEquation2D
    from hit_solver_amd.hit2d import HIT2DConfig, create_grid, initialize_hit_2d

    cfg = HIT2DConfig(nx=nx, ny=ny, target_mach=0.25, initial_kmin=3, initial_kmax=5)
    dom, _, _ = create_grid(cfg)
    eq = EulerEquation2D(gamma=1.4)
    q = initialize_hit_2d(cfg, dom, eq, np.random.default_rng(seed), dtype=dtype)
    return cfg, dom, eq, eq.enforce_physical_state(q)


def test_hit_initial_state_matches_reference(golden):
    print("[init] HIT 32x24 vs golden F6", flush=True)
    _c, _d, _e, q = _hit(32, 24)
    ref = golden["F6_q_32x24"]
    got = q.cpu().numpy()
    for k in range(4):
        e = rel_l2(got[k], ref[k])
        assert e <= 1e-13, f"field {k}: rel L2 {e:.3e}"


def test_hit_initial_512_diagnostics_match_reference(golden):
    from hit_solver_amd.hit2d import compute_diagnostics
    from hit_solver_amd.spatial_operator import PeriodicEulerShockSensor2D

    print("[init] HIT 512^2 initial diagnostics vs golden F6", flush=True)
    _c, dom, eq, q = _hit(512, 512)
    mask = PeriodicEulerShockSensor2D(dom, eq, width=4, jump_threshold=0.04).detect_u8(q)
    count = int(mask.sum().item())
    d = compute_diagnostics(q, eq, dom, weno_fraction=count / 512**2)
    ke, mt, weno, vort, div = (float(x) for x in golden["F6_512_scalars"])
    assert d["kinetic_energy"] == pytest.approx(ke, rel=1e-12)
    assert d["turbulent_mach"] == pytest.approx(mt, rel=1e-12)
    assert d["vorticity_rms"] == pytest.approx(vort, rel=1e-12)
    assert abs(d["divergence_rms"] - div) <= 1e-12
    ref_count = int(golden["F6_512_mask_count"])
    # the device initial state agrees with the reference's to 1e-13, and no
    # node sits that close to a sensor threshold: the count is the same
    assert count == ref_count, f"mask count {count} vs reference {ref_count}"
    print(f"[init] mask count {count} (reference {ref_count})", flush=True)


def test_hit_fp32_initial_state_is_fp64_rounded():
    print("[init] HIT fp32 = rounded fp64", flush=True)
    from hit_solver_amd.equations import EulerEquation2D
    from hit_solver_amd.hit2d import HIT2DConfig, create_grid, initialize_hit_2d

    cfg = HIT2DConfig(nx=48, ny=40, target_mach=0.3, initial_kmin=1, initial_kmax=3)
    dom, _, _ = create_grid(cfg)
    eq = EulerEquation2D()
    a = initialize_hit_2d(cfg, dom, eq, np.random.default_rng(5))
    b = initialize_hit_2d(cfg, dom, eq, np.random.default_rng(5), dtype=torch.float32)
    assert torch.equal(a.float(), b)


def test_hit_initial_state_errors():
    from hit_solver_amd.equations import EulerEquation2D
    from hit_solver_amd.hit2d import HIT2DConfig, create_grid, initialize_hit_2d

    cfg = HIT2DConfig(nx=16, ny=16, initial_kmin=4, initial_kmax=3)  # empty shell
    dom, _, _ = create_grid(cfg)
    with pytest.raises(ValueError, match="zero velocity"):
        initialize_hit_2d(cfg, dom, EulerEquation2D(), np.random.default_rng(0))


@pytest.mark.parametrize("name,factory,dtype", [
    ("c3", OR.config3, "fp64"), ("c308", OR.config3_08, "fp64"), ("c6", OR.config6, "fp64"),
    ("c3_fp32", OR.config3, "fp32"), ("c6_fp32", OR.config6, "fp32"),
])
def test_riemann_initial_states_exact(name, factory, dtype):
    import hit_solver_amd.riemann as R

    print(f"[init] Riemann {name}", flush=True)
    cls = {"c3": R.RiemannConfig3, "c308": R.RiemannConfig3Offset08, "c6": R.RiemannConfig6}[name.split("_")[0]]
    cfg = cls(nx=64, ny=48, dtype=dtype)
    q = cfg.initial_state()
    ref = OR.initial_state(factory(nx=64, ny=48))
    if dtype == "fp32":
        ref = ref.astype(np.float32)
    np.testing.assert_array_equal(q.cpu().numpy(), ref)
