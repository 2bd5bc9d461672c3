"""Serving loop: continuous batching of forced-HIT and Riemann solves on one
GPU, admitted by HBM budget, with device-resident state.

Requests arrive over time (a Poisson stream in the benchmark).  Admitted
solves are grouped by element type -- forced HIT and the Riemann problems
share one batch per dtype, each batch on its own HIP stream so the two
classes' work overlaps on the device.  The engine advances every batch by a round of steps
(a captured HIP graph), reads the control scalars once per round, retires the
solves that finished and admits waiting ones as HBM frees up.  When a class's
membership changes its batch is rebuilt and the running solves move over
(state, control scalars, OU coefficients, noise stream: copy_from).

Launch shapes are pinned per instance (each grid's own reduction tiles, the
wave form of the line sweep where the line allows it, the column-walk form
of the drag stage sums at every batch size), so a solve's
arithmetic does not depend on what else is in the mix: a result is the same
alone and in any mix, bit for bit.

The per-request drivers are the reference's loops: OOP/hit2d.py:764-918 for
forced HIT, OOP/problems/riemann_config3.py:681-749 for the Riemann problems
(configurations 3, 3 with the split at 0.8, 6).
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .batch import BatchRunner, hit_spec_kpar, riemann_spec_kpar
from .hit2d import HIT2DConfig
from .riemann import RiemannConfig3, RiemannConfig3Offset08, RiemannConfig6
from .runtime import make_desc

KINDS = ("c3", "c3_08", "c6", "hit")
_RIEMANN = {"c3": RiemannConfig3, "c3_08": RiemannConfig3Offset08, "c6": RiemannConfig6}


@dataclass
class Request:
    rid: int
    kind: str        # "c3", "c3_08", "c6" or "hit"
    n: int           # grid nx = ny
    dtype: str       # "fp32" or "fp64"
    steps: int       # length of the solve in steps
    arrival: float   # seconds after the stream starts
    seed: int = 1234  # HIT initial-condition seed (forcing stream: seed + 1)


@dataclass
class Result:
    request: Request
    admitted: float = float("nan")
    finished: float = float("nan")
    steps_done: int = 0
    state: np.ndarray | None = None

    @property
    def latency(self):
        return self.finished - self.request.arrival


def riemann_dt0(cfg) -> float:
    """dt of the first step from the quadrant states (piecewise constant initial
    state): cfl / max((|u|+c)/dx + (|v|+c)/dy) (OOP/spatial_operator.py:369-377)."""
    d = cfg.domain
    best = 0.0
    for rho, u, v, p in cfg.quadrants:
        c = np.sqrt(cfg.gamma * p / rho)
        best = max(best, (abs(u) + c) / d.dx + (abs(v) + c) / d.dy)
    return cfg.cfl / best


def request_config(r: Request):
    """Driver configuration of a request: production forced HIT, or the
    Riemann configuration with tfinal sized to r.steps fixed-dt steps."""
    if r.kind not in KINDS:
        raise ValueError(f"unknown request kind {r.kind!r}")
    if r.dtype not in ("fp32", "fp64"):
        raise ValueError("dtype must be fp32 or fp64")
    if r.kind == "hit":
        return HIT2DConfig.production_512(nx=r.n, ny=r.n, dtype=r.dtype, tfinal=1e9, forcing_seed=r.seed,
                                          max_steps=r.steps)
    base = _RIEMANN[r.kind](nx=r.n, ny=r.n, dtype=r.dtype, shear_threshold=None, progress_every=10**9)
    return _RIEMANN[r.kind](nx=r.n, ny=r.n, dtype=r.dtype, shear_threshold=None, progress_every=10**9,
                            tfinal=r.steps * riemann_dt0(base))


def class_key(cfg):
    """Kernel class: what one batch's instances must share (the element
    type; HIT and Riemann solves, schemes and drag settings mix)."""
    return (cfg.torch_dtype,)


_FOOTPRINT: dict = {}


def footprint(cfg) -> int:
    """Device bytes one instance of cfg needs in a batch workspace."""
    key = (type(cfg).__name__, cfg.nx, cfg.ny, cfg.dtype)
    if key not in _FOOTPRINT:
        spec = hit_spec_kpar(cfg)[0] if isinstance(cfg, HIT2DConfig) else riemann_spec_kpar(cfg)[0]
        d, keep = make_desc(spec)
        descs = (_lib.HitxDesc * 1)()
        descs[0] = d
        kp = np.zeros((1, 24))
        h = ctypes.c_void_p()
        check(lib().hitx_batch_create(descs, kp.ctypes.data_as(ctypes.c_void_p), 1, ctypes.byref(h)))
        _FOOTPRINT[key] = int(lib().hitx_batch_workspace_bytes(h))
        lib().hitx_batch_destroy(h)
        del keep
    return _FOOTPRINT[key]


@dataclass
class _Member:
    result: Result
    cfg: object
    fp: int


@dataclass
class _Group:
    members: list = field(default_factory=list)
    runner: BatchRunner | None = None
    slots: list = field(default_factory=list)  # member of each batch slot
    stream: torch.cuda.Stream | None = None
    dirty: bool = True


def poisson_stream(count, rate, seed=0, sizes=tuple(range(128, 1025, 64)), steps=(200, 5000), kinds=KINDS,
                   dtypes=("fp32", "fp64")):
    """count requests with exponential inter-arrival times (rate per second),
    kind, grid, dtype and length drawn uniformly."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        out.append(Request(rid=i, kind=str(rng.choice(kinds)), n=int(rng.choice(sizes)), dtype=str(rng.choice(dtypes)),
                           steps=int(rng.integers(steps[0], steps[1] + 1)), arrival=t, seed=int(rng.integers(0, 2**31))))
    return out


class ServingEngine:
    """Continuous batching of solves under an HBM budget (bytes of batch
    workspace; default: 80% of the device memory free at construction)."""

    def __init__(self, hbm_budget: int | None = None, round_steps: int = 10, max_instances: int = 512,
                 keep_states: bool = False):
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.budget = int(hbm_budget)
        self.round_steps = int(round_steps)
        self.max_instances = int(max_instances)
        self.keep_states = keep_states
        self.groups: dict = {}
        self.used = 0
        self.log: list = []
        # host wall time by phase: rebuilds (capture included), rounds (graph
        # replay to the synchronisation), control reads, idle waits
        self.times = {"rebuilds": 0, "rebuild_s": 0.0, "capture_s": 0.0, "round_s": 0.0, "control_s": 0.0,
                      "idle_s": 0.0}

    # ---------------------------------------------------------------- batches
    def _rebuild(self, key, g: _Group):
        """Rebuild a class's batch after its membership changed.  The running
        solves' state (state, control scalars, OU coefficients, noise stream)
        is staged in small copies and the old batch workspace is freed before
        the new one is allocated, so a rebuild never holds two workspaces."""
        t_start = time.perf_counter()
        old, old_slots = g.runner, g.slots
        g.runner, g.slots = None, []
        carry = {}
        if old is not None:
            torch.cuda.synchronize()
            where = {id(m): j for j, m in enumerate(old_slots)}
            for m in g.members:
                if id(m) in where:
                    carry[id(m)] = old.export_instance(where[id(m)])
            old._sync()
            del old
        if g.members:
            if g.stream is None:
                g.stream = torch.cuda.Stream()
            cfgs = [m.cfg for m in g.members]
            init = [id(m) not in carry for m in g.members]
            caps = [m.result.request.steps for m in g.members]
            torch.cuda.synchronize()
            r = BatchRunner(cfgs, max_steps=caps, init=init, stream=g.stream)
            for i, m in enumerate(g.members):
                if not init[i]:
                    r.import_instance(i, carry.pop(id(m)))
            # pinned launch shapes: per-instance tiles, wave form where the line
            # allows, drag stage sums by column walk whatever the class's size
            r.tune(0, 0, 2, 0)
            r.set_stage_form(2)
            t_cap = time.perf_counter()
            r.capture(self.round_steps)
            self.times["capture_s"] += time.perf_counter() - t_cap
            g.runner, g.slots = r, list(g.members)
        g.dirty = False
        self.times["rebuild_s"] += time.perf_counter() - t_start  # staging, build, initial states, capture
        self.times["rebuilds"] += 1

    # ---------------------------------------------------------------- loop
    def run(self, requests, time_limit: float | None = None):
        """Serve the requests (arrival times relative to the start of the call).
        Returns (results in request order, summary dict)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        results = {r.rid: Result(r) for r in requests}
        queue: list = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        cell_updates = 0
        rounds = 0
        while pending or queue or any(g.members for g in self.groups.values()):
            if time_limit is not None and now() > time_limit:
                break
            t = now()
            while pending and pending[0].arrival <= t:
                queue.append(pending.pop(0))
            # admit in arrival order while the budget allows
            while queue:
                r = queue[0]
                cfg = request_config(r)
                fp = footprint(cfg)
                if fp > self.budget:
                    raise ValueError(f"request {r.rid} needs {fp} bytes, more than the whole budget")
                n_inst = sum(len(g.members) for g in self.groups.values())
                if self.used + fp > self.budget or n_inst >= self.max_instances:
                    break
                queue.pop(0)
                key = class_key(cfg)
                g = self.groups.setdefault(key, _Group())
                g.members.append(_Member(results[r.rid], cfg, fp))
                g.dirty = True
                self.used += fp
                results[r.rid].admitted = now()
            for key, g in self.groups.items():
                if g.dirty:
                    self._rebuild(key, g)
            live = [g for g in self.groups.values() if g.runner is not None]
            if not live:
                if pending:
                    t_idle = time.perf_counter()
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                    self.times["idle_s"] += time.perf_counter() - t_idle
                continue
            t_round = time.perf_counter()
            for g in live:  # one round per class, each on its own stream
                g.runner.replay(1)
            torch.cuda.synchronize()
            self.times["round_s"] += time.perf_counter() - t_round
            rounds += 1
            t = now()
            t_ctl = time.perf_counter()
            for key, g in self.groups.items():
                if g.runner is None:
                    continue
                ctl = g.runner.control()
                done = []
                for i, m in enumerate(g.slots):
                    steps = int(ctl[i, 4])
                    cell_updates += (steps - m.result.steps_done) * m.cfg.nx * m.cfg.ny
                    m.result.steps_done = steps
                    if ctl[i, 0] == 0.0:
                        m.result.finished = t
                        if self.keep_states:
                            m.result.state = g.runner.q[i].detach().cpu().numpy()
                        done.append(m)
                if done:
                    for m in done:
                        g.members.remove(m)
                        self.used -= m.fp
                    g.dirty = True
            for key, g in self.groups.items():
                if g.dirty and not g.members:
                    g.runner, g.slots, g.dirty = None, [], False
            self.times["control_s"] += time.perf_counter() - t_ctl
        wall = now()
        res = [results[r.rid] for r in requests]
        lat = np.array([x.latency for x in res if np.isfinite(x.finished)])
        summary = {
            "requests": len(res), "completed": int(lat.size), "wall_s": round(wall, 3), "rounds": rounds,
            "round_steps": self.round_steps,
            "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
            "mcell_updates_per_s": cell_updates / wall / 1e6 if wall > 0 else None,
            "hbm_budget_bytes": self.budget,
            "phases": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in self.times.items()},
        }
        return res, summary
