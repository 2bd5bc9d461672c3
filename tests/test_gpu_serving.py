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


def _config5_mix():
    """Config-5 sizes: a 1024^2 and a 128^2 solve of every kind (both
    dtypes), plus three 1024^2 FP32 HIT solves so the FP32 HIT class holds
    more than 4096 sensor tiles (6144 + 16)."""
    from hit_solver_amd.serving import KINDS, Request

    reqs, rid = [], 0
    for j, kind in enumerate(KINDS):
        reqs.append(Request(rid, kind, 1024, ("fp64", "fp32")[j % 2], 200, 0.0, seed=40 + j))
        reqs.append(Request(rid + 1, kind, 128, ("fp32", "fp64")[j % 2], 260, 0.005 * j, seed=50 + j))
        rid += 2
    for k in range(3):
        reqs.append(Request(rid + k, "hit", 1024, "fp32", 200, 0.02, seed=60 + k))
    return reqs


def test_serving_config5_sizes_same_alone_and_in_mix():
    """1024^2 and 128^2 solves of every kind in one stream: each result is
    bit-identical to the solve served alone, including HIT solves in a class
    above 4096 sensor tiles (the drag stage sums keep one form)."""
    from hit_solver_amd.serving import ServingEngine

    reqs = _config5_mix()
    print(f"[serving] config-5 mix: {len(reqs)} requests", flush=True)
    mixed, summary = ServingEngine(round_steps=10, keep_states=True).run(reqs)
    assert summary["completed"] == len(reqs)
    for res in mixed:
        r = res.request
        assert res.steps_done == r.steps, (r.rid, res.steps_done)
        print(f"[serving] alone: request {r.rid} {r.kind} {r.n}^2 {r.dtype}", flush=True)
        alone, _ = ServingEngine(round_steps=10, keep_states=True).run([r])
        np.testing.assert_array_equal(res.state, alone[0].state, err_msg=f"request {r.rid}")


def test_serving_fp64_request_vs_oracle():
    """A served FP64 forced-HIT request against the NumPy FP64 oracle run of
    the same configuration (production closure, 128^2, 20 steps)."""
    import dataclasses

    from hit_solver_amd.serving import Request, ServingEngine, request_config
    from oracle import hit as OH

    r = Request(0, "hit", 128, "fp64", 20, 0.0, seed=77)
    res, summary = ServingEngine(round_steps=10, keep_states=True).run([r])
    assert summary["completed"] == 1 and res[0].steps_done == 20
    cfg = request_config(r)
    names = {f.name for f in dataclasses.fields(OH.HITParams)}
    ref = OH.HITRun(OH.HITParams(**{k: v for k, v in dataclasses.asdict(cfg).items() if k in names}))
    ref.run()
    assert ref.steps == 20
    for k in range(4):
        e = float(np.linalg.norm(res[0].state[k] - ref.q[k]) / np.linalg.norm(ref.q[k]))
        print(f"[serving] FP64 request vs oracle field {k}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-10, (k, e)


def test_serving_rebuild_holds_one_workspace():
    """A class whose batch is more than half the budget grows by one solve:
    the rebuild stages the running solve's state and frees the old workspace
    before allocating the new one, so device memory never holds both."""
    from hit_solver_amd.serving import Request, ServingEngine, footprint, request_config

    big = Request(0, "hit", 1024, "fp64", 100, 0.0, seed=5)
    small = Request(1, "hit", 128, "fp64", 10, 0.05, seed=6)
    fb, fs = footprint(request_config(big)), footprint(request_config(small))
    budget = fb + fs + (1 << 20)
    assert fb > budget // 2
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    res, summary = ServingEngine(hbm_budget=budget, round_steps=5, keep_states=False).run([big, small])
    assert summary["completed"] == 2
    peak = torch.cuda.max_memory_allocated() - base
    staged = 4 * 1024 * 1024 * 8 + (1 << 20)  # the running solve's state and its small arrays
    print(f"[serving] peak {peak / 2**20:.1f} MiB, budget {budget / 2**20:.1f} MiB", flush=True)
    assert peak <= budget + staged + (64 << 20), (peak, budget)
