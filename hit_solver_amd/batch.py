"""Batched runs: many HIT and Riemann instances of different grid sizes
advance together through ``hitx_batch_step`` (one launch per kernel kind per
stage for all instances -- two for the line sweeps when periodic and
non-periodic instances are mixed -- with per-instance scalar control on the
device).

Each instance follows the reference's step loop for its own configuration
(OOP/hit2d.py:764-918 or OOP/problems/riemann_config3.py:681-749).  The
forcing noise of each HIT instance is its own NumPy PCG64 stream
(default_rng(forcing_seed + 1)), generated on the device inside the step, so a
step needs no host work and no synchronisation.
"""

from __future__ import annotations

import contextlib
import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .forcing import shell_modes
from .hit2d import HIT2DConfig, _make_equation, create_grid, drag_modes, initialize_hit_2d
from .riemann import RiemannConfig3, hyperviscosity_enabled_for_config
from .rng import pcg64_words, pcg64_words_rewind
from .runtime import PlanSpec, get_plan, make_desc
from .spatial_operator import PeriodicHyperviscosity2D

NAN = float("nan")
# control-scalar slots (include/hitx.h)
C_ACTIVE, C_TIME, C_TFINAL, C_DT, C_STEP, C_MAXSTEPS = 0, 1, 2, 3, 4, 5
C_HVAPPS, C_POWER, C_ALPHA, C_NSTEPS = 7, 11, 12, 26
C_NPRE, C_NPREW = 27, 28  # next step's noise drawn during the last step; the PCG64 words it consumed
C_FXX, C_FYY, C_FXY, C_PIN, C_PBEFORE, C_ATGT, C_MACH, C_PDES = 18, 19, 20, 21, 22, 23, 24, 25
BUF_STATE, BUF_NOISE, BUF_CTL, BUF_PHI, BUF_RNG = 0, 1, 2, 3, 4


def _none(v):
    return NAN if v is None else float(v)


def hit_spec_kpar(c: HIT2DConfig):
    """PlanSpec and run parameters of one forced-HIT instance."""
    domain, _, _ = create_grid(c)
    eq = _make_equation(c)
    hv = PeriodicHyperviscosity2D(domain, mn=c.hyperviscosity_mn)
    dm = drag_modes(domain, c.large_scale_drag_kmax) if c.large_scale_drag > 0.0 and c.large_scale_drag_kmax > 0.0 \
        else np.zeros((0, 2), int)
    fmodes, _, _ = shell_modes(domain, 0.0 if c.forcing_kmin is None else c.forcing_kmin, c.forcing_kmax)
    spec = PlanSpec(
        nx=c.nx, ny=c.ny, dx=domain.dx, dy=domain.dy, periodic=True, dtype=c.torch_dtype, gamma=c.gamma,
        sensor_width=c.sensor_width, compression_threshold=c.compression_threshold,
        jump_threshold=c.jump_threshold, shear_threshold=c.shear_threshold, viscosity=c.viscosity,
        prandtl=c.prandtl, hv_coef=hv.coefficients(), forcing=True,
        drag_coef=c.large_scale_drag if len(dm) else 0.0, drag_modes=tuple(int(v) for v in dm.reshape(-1)),
        cooling_time=c.cooling_time if (c.cooling_time is not None and c.cooling_time > 0.0) else None,
        cooling_p_ref=c.cooling_target_pressure, force_modes=tuple(int(v) for v in fmodes.reshape(-1)),
    )
    h = min(domain.dx, domain.dy)
    tgt = c.target_mach if c.mach_control_target is None else c.mach_control_target
    k = [0.0] * _lib_kpar()
    k[0] = c.cfl
    k[1] = 0.25 * h**2 / c.viscosity if c.viscosity > 0.0 else NAN
    k[2] = 0.0
    k[3] = float(c.hyperviscosity_interval)
    k[4] = 1.0 if c.hyperviscosity_mn > 0.0 else 0.0
    k[5] = 1.0 if c.mach_control else 0.0
    k[6], k[7], k[8] = float(tgt), c.mach_control_memory, c.mach_control_exponent
    k[9], k[10] = _none(c.mach_control_min_power), _none(c.mach_control_max_power)
    k[11], k[12], k[13] = _none(c.target_energy_injection), c.min_forcing_power, _none(c.max_forcing_rescale)
    k[14], k[15], k[16], k[17] = c.forcing_alpha_memory, c.forcing_rms, c.forcing_correlation_time, 1.0
    k[18] = c.tfinal
    k[19] = 1e300 if c.max_steps is None else float(c.max_steps)
    return spec, k, domain, eq


def riemann_spec_kpar(c: RiemannConfig3, max_steps=None):
    from .riemann import LocalHyperviscosity2D

    d = c.domain
    hv = LocalHyperviscosity2D(d, mn=c.mn, density_weight=c.hyperviscosity_density_weight,
                               momentum_weight=c.hyperviscosity_momentum_weight,
                               energy_weight=c.hyperviscosity_energy_weight)
    spec = PlanSpec(
        nx=c.nx, ny=c.ny, dx=d.dx, dy=d.dy, periodic=False, dtype=c.torch_dtype, scheme=c.scheme, gamma=c.gamma,
        sensor_width=c.sensor_width, boundary_guard=c.boundary_guard,
        compression_threshold=c.compression_threshold, jump_threshold=c.jump_threshold,
        shear_threshold=c.shear_threshold, hv_coef=hv.coefficients(), guard_cells=c.guard_cells,
    )
    k = [0.0] * _lib_kpar()
    k[0] = c.cfl
    k[1] = NAN
    k[2] = 1.0
    k[3] = float(c.hyperviscosity_interval)
    k[4] = 1.0 if hyperviscosity_enabled_for_config(c) else 0.0
    k[18] = c.tfinal
    k[19] = 1e300 if max_steps is None else float(max_steps)
    return spec, k


def hit_batched_configs(dtype="fp32"):
    """The batched forced-HIT workload: 64 instances, 16 each at 128^2, 256^2,
    384^2 and 512^2, seeds 0-63, Mach (0.1, 0.25, 0.4)[seed % 3] (also the
    Mach-control target), production settings otherwise."""
    out, seed = [], 0
    for n in (128, 256, 384, 512):
        for _ in range(16):
            mt = (0.1, 0.25, 0.4)[seed % 3]
            out.append(HIT2DConfig.production_512(nx=n, ny=n, forcing_seed=seed, target_mach=mt,
                                                  mach_control_target=mt, dtype=dtype, tfinal=1e9))
            seed += 1
    return out


def _lib_kpar():
    return 24  # HITX_KPAR


class BatchRunner:
    """All instances share the dtype (one kernel instantiation per batch);
    forced HIT and Riemann instances, schemes, sensor widths and drag settings
    may be mixed.  Each instance runs the same reductions and sweeps as in a
    batch of its own kind, so mixing does not change its results.

    max_steps: None, one cap for every instance, or a list (one per instance).
    init: True, or a list of booleans: instances left uninitialised are to be
    filled by copy_from (moving a running instance between batches).
    stream: the torch stream every launch of this batch goes to (default: the
    current stream at each call)."""

    def __init__(self, configs, max_steps=None, init=True, stream=None):
        if not configs:
            raise ValueError("empty batch")
        self.configs = list(configs)
        n = len(self.configs)
        self.stream = stream
        self.is_hit = [isinstance(c, HIT2DConfig) for c in self.configs]
        self.hit = all(self.is_hit)  # every instance forced HIT
        caps = list(max_steps) if isinstance(max_steps, (list, tuple)) else [max_steps] * n
        todo = list(init) if isinstance(init, (list, tuple)) else [bool(init)] * n
        if len(caps) != n or len(todo) != n:
            raise ValueError("max_steps and init lists need one entry per instance")
        specs, kpars, self.domains = [], [], []
        for c, cap, is_hit in zip(self.configs, caps, self.is_hit):
            if not is_hit and not isinstance(c, RiemannConfig3):
                raise ValueError(f"batch instances are HIT2DConfig or RiemannConfig3, not {type(c).__name__}")
            if is_hit:
                if cap is not None:
                    c = dataclasses.replace(c, max_steps=int(cap))
                spec, k, domain, _ = hit_spec_kpar(c)
                self.domains.append(domain)
            else:
                spec, k = riemann_spec_kpar(c, cap)
                self.domains.append(c.domain)
            specs.append(spec)
            kpars.append(k)
        self.specs, self.kpars = specs, kpars
        descs = (_lib.HitxDesc * n)()
        keep = []
        for i, sp in enumerate(specs):
            d, kp = make_desc(sp)
            descs[i] = d
            keep.append(kp)
        kp = np.ascontiguousarray(np.array(kpars, dtype=np.float64))
        h = ctypes.c_void_p()
        check(lib().hitx_batch_create(descs, kp.ctypes.data_as(ctypes.c_void_p), n, ctypes.byref(h)))
        self.handle = h
        self.n = n
        nbytes = int(lib().hitx_batch_workspace_bytes(h))
        with self._ctx():
            self.ws = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
        self._sync()
        check(lib().hitx_batch_bind(h, ctypes.c_void_p(self.ws.data_ptr()), nbytes))
        self.q = [self._view(i, BUF_STATE, sp.dtype, (4, sp.ny, sp.nx)) for i, sp in enumerate(specs)]
        self.ctl_views = [self._view(i, BUF_CTL, torch.float64, (32,)) for i in range(n)]
        # forcing buffers of the HIT instances (None for Riemann instances)
        self.noise = [self._view(i, BUF_NOISE, torch.float64, (sp.ny, sp.nx)) if h else None
                      for i, (sp, h) in enumerate(zip(specs, self.is_hit))]
        self.rng_state = [self._view(i, BUF_RNG, torch.int64, (4,)) if h else None for i, h in enumerate(self.is_hit)]
        self.phi = [self._view(i, BUF_PHI, torch.float64, (max(1, len(sp.force_modes) // 2), 2)) if h else None
                    for i, (sp, h) in enumerate(zip(specs, self.is_hit))]
        with self._ctx():
            for i in range(n):
                if todo[i]:
                    self._init_instance(i)
            if not self.hit:
                self.fixed_dt()
        self._graph = None
        self._sync()

    def _ctx(self):
        return torch.cuda.stream(self.stream) if self.stream is not None else contextlib.nullcontext()

    def _sync(self):
        if self.stream is not None:
            self.stream.synchronize()
        else:
            torch.cuda.synchronize()

    def _init_instance(self, i):
        """Initial state (enforce, + outflow guard), control scalars and noise
        stream of instance i, as the single-instance drivers set them up."""
        c, sp, k, hit = self.configs[i], self.specs[i], self.kpars[i], self.is_hit[i]
        if hit:
            domain = self.domains[i]
            q0 = initialize_hit_2d(c, domain, _make_equation(c), np.random.default_rng(c.forcing_seed),
                                   dtype=sp.dtype)
        else:
            q0 = c.initial_state()
        get_plan(sp).enforce(q0, apply_guard=not hit, out=self.q[i])
        ctl = torch.zeros(32, dtype=torch.float64)
        ctl[C_ACTIVE] = 1.0
        ctl[C_TFINAL] = c.tfinal
        ctl[C_MAXSTEPS] = k[19]
        ctl[C_POWER] = k[11]
        ctl[C_ALPHA] = 1.0
        self.ctl_views[i].copy_(ctl)
        if hit:
            w = pcg64_words(np.random.default_rng(c.forcing_seed + 1).bit_generator)
            self.rng_state[i].copy_(torch.from_numpy(w.view(np.int64).copy()))

    def copy_from(self, i, other: "BatchRunner", j):
        """Continue other's instance j as this batch's instance i (same
        configuration): state, control scalars, OU coefficients, noise stream."""
        self.import_instance(i, other.export_instance(j))

    def export_instance(self, j):
        """Copies of what continues instance j elsewhere (state, control
        scalars, OU coefficients, noise stream): a few arrays of the state's
        size, so the batch workspace can be freed before the next batch is
        built.  A step draws the next step's noise while it runs; that draw
        is not carried over: the noise stream is moved back by the words it
        consumed, so the next batch draws the same values again."""
        with self._ctx():
            out = {"q": self.q[j].clone(), "ctl": self.ctl_views[j].clone()}
            if self.is_hit[j]:
                out["phi"] = self.phi[j].clone()
                out["rng"] = self.rng_state[j].clone()
        if self.is_hit[j]:
            self._sync()
            ctl = out["ctl"].cpu().numpy()
            if ctl[C_NPRE] != 0.0:
                w = pcg64_words_rewind(out["rng"].cpu().numpy(), int(ctl[C_NPREW]))
                out["rng"].copy_(torch.from_numpy(w.view(np.int64).copy()))
                out["ctl"][C_NPRE] = 0.0
                out["ctl"][C_NPREW] = 0.0
        return out

    def import_instance(self, i, saved):
        """Continue an exported instance as this batch's instance i (same
        configuration)."""
        with self._ctx():
            self.q[i].copy_(saved["q"])
            self.ctl_views[i].copy_(saved["ctl"])
            if self.is_hit[i]:
                self.phi[i].copy_(saved["phi"])
                self.rng_state[i].copy_(saved["rng"])

    def fixed_dt(self):
        """Fixed dt of the Riemann instances that have none yet."""
        check(lib().hitx_batch_fixed_dt(self.handle, self._stream()))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().hitx_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def _stream(self):
        st = self.stream if self.stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(st.cuda_stream)

    def _view(self, i, which, dtype, shape):
        p = ctypes.c_void_p()
        check(lib().hitx_batch_buffer(self.handle, i, which, ctypes.byref(p)))
        off = p.value - self.ws.data_ptr()
        nb = int(np.prod(shape)) * torch.tensor([], dtype=dtype).element_size()
        return self.ws[off:off + nb].view(dtype).view(*shape)

    def step(self):
        """One step of every active instance (asynchronous, no host work)."""
        check(lib().hitx_batch_step(self.handle, self._stream()))

    def capture(self, steps: int = 1):
        """Record `steps` batch steps as one HIP graph (the step has no host
        work: dt, control and noise are on the device).  Replaying it advances
        every active instance; finished instances idle through it."""
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        keep, self.stream = self.stream, None
        try:
            with torch.cuda.graph(g, stream=s):
                for _ in range(steps):
                    self.step()
        finally:
            self.stream = keep
        torch.cuda.synchronize()
        self._graph, self._graph_steps = g, steps
        return g

    def replay(self, n: int = 1):
        """Replay the captured graph n times (n * captured steps), on this
        batch's stream."""
        if getattr(self, "_graph", None) is None:
            raise RuntimeError("capture() first")
        with self._ctx():
            for _ in range(n):
                self._graph.replay()

    def tune(self, rb: int, rbp: int, line_form: int = 0, sensor_form: int = 0):
        """Launch shapes of every instance: rows per wave of the stage
        reductions and of the forcing projection, the line-sweep form (0 auto,
        1 block per line, 2 wavefront per line, 3 wavefront per line with the
        y-lines read from the state through an LDS slab instead of a
        transposed copy, 4 wavefront per line with the x-sweep and the stage
        update of each row panel in one launch where the batch allows it:
        FP32, periodic grids, both line lengths in the wave form; otherwise
        as 2) and the sensor form (0 auto,
        1 fused tile, 2 split with the column-walk raw mask, 3 split with the
        tile raw mask)."""
        torch.cuda.synchronize()
        check(lib().hitx_batch_tune(self.handle, int(rb), int(rbp), int(line_form), int(sensor_form)))
        self._graph = None  # grid sizes changed

    def forms(self):
        """Launch forms in effect: dict of booleans (x-sweep fused into the
        update, x- and y-lines in the wave form, slab y-sweep, split sensor,
        drag stage sums by column walk)."""
        out = (ctypes.c_int32 * 6)()
        check(lib().hitx_batch_forms(self.handle, ctypes.cast(out, ctypes.c_void_p)))
        keys = ("xfused", "wave_x", "wave_y", "slab_y", "split_sensor", "stage_walk")
        return {k: bool(v) for k, v in zip(keys, out)}

    def serial_draws(self, on: bool = True):
        """Make every noise draw take the serial path (a check of that path:
        the same values, projected from the noise buffer)."""
        check(lib().hitx_batch_serial_draws(self.handle, 1 if on else 0))
        self._graph = None

    def set_stage_form(self, form: int):
        """Form of the drag stage reductions: 0 auto (tiles while the batch has
        at most 4096 sensor tiles, column walk above), 1 tiles, 2 column walk.
        Fixing it keeps an instance's summation order independent of the
        rest of the batch."""
        torch.cuda.synchronize()
        check(lib().hitx_batch_stage_form(self.handle, int(form)))
        self._graph = None

    def timing(self, on: bool):
        """HIP events around every stage-update launch of eager steps."""
        check(lib().hitx_batch_timing_enable(self.handle, int(on)))

    def timing_read(self):
        """(summed stage-update milliseconds, launches) since the last read."""
        ms, n = ctypes.c_double(), ctypes.c_int32()
        check(lib().hitx_batch_timing_read(self.handle, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def time_updates(self, reps: int = 5):
        """Average stage-update kernel milliseconds: the three stage updates
        of every instance launched reps times back to back between two events
        (per periodicity present).  The state is not a solution afterwards."""
        ms = ctypes.c_double()
        check(lib().hitx_batch_time_updates(self.handle, int(reps), self._stream(), ctypes.byref(ms)))
        # forced HIT instances are periodic, the Riemann problems not: one
        # update launch per periodicity present
        launches = 3 * int(reps) * (int(any(self.is_hit)) + int(not all(self.is_hit)))
        return ms.value / launches, launches

    def autotune(self, candidates=None, steps: int = 3):
        """Time a few steps per candidate (rb, rbp, line_form, sensor_form)
        on a copy of the batch's device state, restore the state and keep the
        fastest candidate.  Returns (best, {candidate: ms per step})."""
        if candidates is None:
            candidates = [(rb, rbp, lf, sf) for lf in (1, 2, 4) for sf in (1, 2, 3) for rb in (4, 16, 32)
                          for rbp in ((16, 64) if any(self.is_hit) else (16,))]
        torch.cuda.synchronize()
        backup = self.ws.clone()
        times = {}
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for cand in candidates:
            self.tune(*cand)
            self.step()
            e0.record()
            for _ in range(steps):
                self.step()
            e1.record()
            torch.cuda.synchronize()
            times[cand] = e0.elapsed_time(e1) / steps
            self.ws.copy_(backup)
        torch.cuda.synchronize()
        del backup
        best = min(times, key=times.get)
        self.tune(*best)
        return best, times

    def control(self):
        """Control scalars of all instances (host numpy, [n, 32]); synchronises."""
        with self._ctx():
            out = torch.stack(self.ctl_views).cpu().numpy()
        return out

    def active(self):
        return self.control()[:, C_ACTIVE] != 0.0

    def run(self, max_steps=None):
        k = 0
        while self.active().any() and (max_steps is None or k < max_steps):
            self.step()
            k += 1
        return k
