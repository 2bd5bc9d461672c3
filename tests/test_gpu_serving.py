"""Serving loop: continuous batching, HBM-budget admission, results the same
alone and in the mix."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _mix():
    from hit_solver_amd.serving import Request

    return [Request(0, "c3", 128, "fp64", 40, 0.00),
            Request(1, "hit", 64, "fp32", 25, 0.00, seed=11),
            Request(2, "c6", 96, "fp32", 30, 0.01),
            Request(3, "hit", 128, "fp64", 35, 0.02, seed=12),
            Request(4, "c3_08", 64, "fp64", 50, 0.02),
            Request(5, "hit", 64, "fp32", 15, 0.03, seed=13),
            Request(6, "c3", 192, "fp32", 20, 0.04)]


def test_serving_results_same_alone_and_in_mix():
    """Launch shapes are pinned per instance, so a solve's arithmetic does not
    depend on the mix: every result is bit-identical to the solve served alone,
    though in the mix instances join and leave batches while it runs."""
    from hit_solver_amd.serving import ServingEngine

    reqs = _mix()
    mixed, summary = ServingEngine(round_steps=5, keep_states=True).run(reqs)
    assert summary["completed"] == len(reqs)
    for res in mixed:
        r = res.request
        assert res.steps_done == r.steps, (r.rid, res.steps_done)
        alone, _ = ServingEngine(round_steps=5, keep_states=True).run([r])
        np.testing.assert_array_equal(res.state, alone[0].state, err_msg=f"request {r.rid}")


def test_serving_budget_admission():
    """A budget that holds two of the solves at a time: all complete, the
    others wait in arrival order, and results still match the solve alone."""
    from hit_solver_amd.serving import ServingEngine, footprint, request_config

    reqs = _mix()[:5]
    fps = [footprint(request_config(r)) for r in reqs]
    budget = sorted(fps)[-1] + sorted(fps)[-2]
    eng = ServingEngine(hbm_budget=budget, round_steps=5, keep_states=True)
    res, summary = eng.run(reqs)
    assert summary["completed"] == len(reqs)
    assert eng.used == 0
    admitted = sorted(res, key=lambda x: x.admitted)
    assert [x.request.rid for x in admitted] == [r.rid for r in sorted(reqs, key=lambda r: r.arrival)]
    r = reqs[3]
    alone, _ = ServingEngine(round_steps=5, keep_states=True).run([r])
    np.testing.assert_array_equal(res[3].state, alone[0].state)
