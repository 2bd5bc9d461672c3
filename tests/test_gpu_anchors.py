"""Full-size device runs against the reference's own published run results.

The reference ships five 512x512 Riemann runs (CuPy backend, FP64) with their
configuration and end-of-run diagnostics (results/riemann_config3/*,
results/riemann_config3_08/*, results/riemann_config6/*; copied unchanged
into tests/golden/reference_run_anchors.json by make_anchors.py).  The device
runs the same configuration from the same initial state.

What must agree exactly: the fixed-dt step count, the number of
hyperviscosity passes, the final time and the state flags (finite, positive
density and pressure).

What can agree only loosely: these 512x512 flows are sensitive to rounding.
Started from an initial state perturbed by one ulp per entry, the device run
moves 1e-4 (relative L2) from itself in 100 steps and 0.5-2% by the end, and
its end-state diagnostics move by up to 0.3% (extrema), 1.5% (minimum
density, configuration 6), 1% (maximum |vorticity|) and 6e-4 (sensor
fraction).  The reference's CuPy runs lie outside that envelope (for
configuration 3: minimum density 0.109 and a pressure at the 1e-10 floor
against 0.137 and 0.028 here, 3% on the density maximum), while the same
device code agrees with the NumPy oracle -- the reference's algorithm -- to
1e-10 after the first 10 steps at this size.  The CuPy results therefore pin
the full-size runs only to a few percent: maxima within 4%, sensor fraction
within 0.015, maximum |vorticity| within 20%.
"""

import json
import math
from pathlib import Path

import pytest
import torch

pytestmark = pytest.mark.gpu

ANCHORS = Path(__file__).resolve().parent / "golden" / "reference_run_anchors.json"
RUNS = json.loads(ANCHORS.read_text())["runs"]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _config(c):
    import hit_solver_amd.riemann as R

    if c["configuration_number"] == 6:
        cls = R.RiemannConfig6
    elif c["x_split"] == 0.8:
        cls = R.RiemannConfig3Offset08
    else:
        cls = R.RiemannConfig3
    keys = ("nx", "ny", "tfinal", "cfl", "gamma", "x_split", "y_split", "scheme", "sensor_width", "jump_threshold",
            "compression_threshold", "shear_threshold", "boundary_guard", "mn", "hyperviscosity_interval",
            "hyperviscosity_density_weight", "hyperviscosity_momentum_weight", "hyperviscosity_energy_weight",
            "guard_cells")
    return cls(progress_every=10**9, **{k: c[k] for k in keys})


@pytest.mark.parametrize("run", RUNS, ids=[r["source"].split("/")[2] for r in RUNS])
def test_riemann_512_matches_reference_results(run):
    from hit_solver_amd.riemann import RiemannSolver

    cfg = _config(run["config"])
    ref = run["diagnostics"]
    print(f"[anchors] {run['source']}: {cfg.nx}x{cfg.ny}, T {cfg.tfinal}, CFL {cfg.cfl}, mn {cfg.mn}", flush=True)
    s = RiemannSolver(cfg)
    assert s.n_steps == ref["steps"]
    s.run()
    d = s.diagnostics()
    got = {k: getattr(d, k) for k in ref if k != "backend"}
    print(f"[anchors] device {got}", flush=True)
    assert d.steps == ref["steps"] and d.hyperviscosity_applications == ref["hyperviscosity_applications"]
    assert d.time == pytest.approx(ref["time"], rel=1e-12)
    for k in ("has_nonfinite", "density_positive", "pressure_positive", "hyperviscosity_enabled", "scheme"):
        assert got[k] == ref[k], k
    for k in ("density_max", "pressure_max"):
        assert got[k] == pytest.approx(ref[k], rel=4e-2), k
    assert got["max_abs_vorticity"] == pytest.approx(ref["max_abs_vorticity"], rel=0.2)
    assert abs(got["final_sensor_fraction"] - ref["final_sensor_fraction"]) <= 0.015
    assert math.isfinite(got["density_min"]) and got["density_min"] > 0.0
