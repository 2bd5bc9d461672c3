"""Plans, workspaces and thin callers of the hitx C ABI.

A :class:`Plan` owns one ``hitx_plan*`` (constant tables only) and a torch
uint8 workspace on the plan's device.  All arrays are torch tensors on the
current HIP device; every call is asynchronous on torch's current stream.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import HitxDesc, check, lib


def require_device(t: torch.Tensor, name: str, dtype=None, shape=None) -> None:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ValueError(f"{name} must be a torch tensor on the HIP device")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


@dataclass(frozen=True)
class PlanSpec:
    """Everything a plan is built from (hashable: plans are cached by spec)."""

    nx: int
    ny: int
    dx: float
    dy: float
    periodic: bool
    dtype: torch.dtype = torch.float64
    scheme: str = "hybrid"
    gamma: float = 1.4
    rho_floor: float = 1e-10
    p_floor: float = 1e-10
    sensor_width: int = 4
    boundary_guard: int = 0
    compression_threshold: float = 2.5
    jump_threshold: float = 0.04
    shear_threshold: float | None = None
    viscosity: float = 0.0
    prandtl: float = 0.72
    gas_constant: float = 1.0
    hv_coef: tuple = (0.0, 0.0, 0.0, 0.0)
    guard_cells: int = 0
    forcing: bool = False
    drag_coef: float = 0.0
    drag_modes: tuple = ()
    cooling_time: float | None = None
    cooling_p_ref: float | None = None
    force_modes: tuple = ()

    def shear_enabled(self) -> bool:
        if self.shear_threshold is None:
            return False
        # periodic sensor: any threshold (OOP/spatial_operator.py:277); the
        # non-periodic one needs > 0 (OOP/problems/riemann_config3.py:303)
        return True if self.periodic else self.shear_threshold > 0.0


def make_desc(spec: PlanSpec):
    """hitx_desc for a spec; returns (desc, arrays that must outlive it)."""
    s = spec
    d = HitxDesc()
    d.abi_version = _lib.HITX_ABI_VERSION
    d.nx, d.ny = s.nx, s.ny
    d.periodic = int(bool(s.periodic))
    d.dtype = _lib.HITX_F64 if s.dtype == torch.float64 else _lib.HITX_F32
    d.scheme = _lib.HITX_WENO if s.scheme == "weno" else _lib.HITX_HYBRID
    d.dx, d.dy = s.dx, s.dy
    d.dx2, d.dy2 = s.dx**2, s.dy**2
    d.gamma, d.rho_floor, d.p_floor = s.gamma, s.rho_floor, s.p_floor
    d.sensor_width = s.sensor_width
    d.boundary_guard = s.boundary_guard
    d.shear_enabled = int(s.shear_enabled())
    d.compression_threshold = s.compression_threshold
    d.jump_threshold = s.jump_threshold
    d.shear_threshold = 0.0 if s.shear_threshold is None else float(s.shear_threshold)
    d.viscosity, d.prandtl, d.gas_constant = s.viscosity, s.prandtl, s.gas_constant
    for k in range(4):
        d.hv_coef[k] = s.hv_coef[k]
    d.guard_cells = s.guard_cells
    d.forcing_enabled = int(s.forcing)
    d.drag_enabled = int(s.drag_coef > 0.0 and len(s.drag_modes) > 0)
    d.drag_coef = s.drag_coef
    d.cooling_enabled = int(s.cooling_time is not None and s.cooling_time > 0.0)
    d.cooling_time = 0.0 if s.cooling_time is None else s.cooling_time
    d.cooling_p_ref = (1.0 / s.gamma) if s.cooling_p_ref is None else s.cooling_p_ref
    dm = np.ascontiguousarray(np.array(s.drag_modes, dtype=np.int32).reshape(-1, 2))
    fm = np.ascontiguousarray(np.array(s.force_modes, dtype=np.int32).reshape(-1, 2))
    d.n_drag_modes, d.n_force_modes = len(dm), len(fm)
    d.drag_modes_host = dm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    d.force_modes_host = fm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    return d, (dm, fm)


class Plan:
    def __init__(self, spec: PlanSpec):
        self.spec = spec
        d, self._keep = make_desc(spec)
        h = ctypes.c_void_p()
        check(lib().hitx_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        self.ws_bytes = int(lib().hitx_workspace_bytes(h))
        self._ws = None
        self.red = torch.zeros(64, dtype=torch.float64, device="cuda")
        self.device = torch.cuda.current_device()

    @property
    def ws(self):
        """The plan's workspace, allocated (zeroed: the reduction ticket starts
        at 0) on the first call that needs it, so plans used only for
        pointwise work (enforce, initial states) hold no grid-sized buffers."""
        if self._ws is None:
            self._ws = torch.zeros(self.ws_bytes, dtype=torch.uint8, device="cuda")
        return self._ws

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().hitx_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    # ---------------------------------------------------------------- checks
    @property
    def shape(self):
        return (4, self.spec.ny, self.spec.nx)

    def _q(self, q, name="q"):
        require_device(q, name, self.spec.dtype, self.shape)

    def _f(self, a, name):
        require_device(a, name, self.spec.dtype, (self.spec.ny, self.spec.nx))

    # ---------------------------------------------------------------- calls
    def sensor(self, q, enforce_first: bool, out=None):
        self._q(q)
        if out is None:
            out = torch.empty((self.spec.ny, self.spec.nx), dtype=torch.uint8, device=q.device)
        check(lib().hitx_sensor(self.handle, _p(q), int(enforce_first), _p(out), _p(self.ws),
                                self.ws_bytes, _stream()))
        return out

    def rhs(self, q, out=None):
        self._q(q)
        out = torch.empty_like(q) if out is None else out
        check(lib().hitx_rhs(self.handle, _p(q), _p(out), _p(self.ws), self.ws_bytes, _stream()))
        return out

    def stage(self, stage, q0, qs, out, dt, fx=None, fy=None):
        self._q(q0, "q0")
        self._q(qs, "qs")
        self._q(out, "out")
        if self.spec.forcing:
            self._f(fx, "fx")
            self._f(fy, "fy")
        check(lib().hitx_rk3_stage(self.handle, int(stage), _p(q0), _p(qs), _p(out), _p(fx), _p(fy),
                                   float(dt), _p(self.ws), self.ws_bytes, _stream()))
        return out

    def hyperviscosity(self, q, mask_u8, apply_guard: bool, out=None, invert: bool = False):
        """Filter where mask != 0 (invert=False: mask marks active nodes) or
        where mask == 0 (invert=True: mask is the WENO-node mask)."""
        self._q(q)
        require_device(mask_u8, "mask", torch.uint8, (self.spec.ny, self.spec.nx))
        out = torch.empty_like(q) if out is None else out
        check(lib().hitx_hyperviscosity(self.handle, _p(q), _p(mask_u8), int(invert), _p(out), int(apply_guard),
                                        _p(self.ws), self.ws_bytes, _stream()))
        return out

    def enforce(self, q, apply_guard: bool, out=None):
        self._q(q)
        out = torch.empty_like(q) if out is None else out
        check(lib().hitx_enforce(self.handle, _p(q), _p(out), int(apply_guard), _stream()))
        return out

    def step_reduce(self, q):
        """(max spectral radius, sum(u^2+v^2), sum(c)) -- synchronises."""
        self._q(q)
        check(lib().hitx_step_reduce(self.handle, _p(q), _p(self.red), _p(self.ws), self.ws_bytes,
                                     _stream()))
        r = self.red[:3].cpu().numpy()
        return float(r[0]), float(r[1]), float(r[2])

    def primitives(self, q):
        self._q(q)
        rho, u, v, p = (torch.empty((self.spec.ny, self.spec.nx), dtype=q.dtype, device=q.device)
                        for _ in range(4))
        check(lib().hitx_primitives(self.handle, _p(q), _p(rho), _p(u), _p(v), _p(p), _stream()))
        return rho, u, v, p

    def viscous_rhs(self, q, out=None):
        self._q(q)
        out = torch.empty_like(q) if out is None else out
        check(lib().hitx_viscous_rhs(self.handle, _p(q), _p(out), _stream()))
        return out

    def diagnostics(self, q, mask_u8=None, lowpass_modes=(), drag_coef: float = 0.0, fields: bool = False):
        """Diagnostic sums of q in double (HITX_DIAG_* slots, synchronises) and,
        when fields, the device planes [6, ny, nx] = rho, u, v, p, divergence,
        vorticity (float64)."""
        self._q(q)
        if mask_u8 is not None:
            require_device(mask_u8, "mask", torch.uint8, (self.spec.ny, self.spec.nx))
        modes = np.ascontiguousarray(np.asarray(lowpass_modes, dtype=np.int32).reshape(-1, 2))
        need = int(lib().hitx_diag_workspace_bytes(self.handle, int(fields)))
        if getattr(self, "_dws", None) is None or self._dws.numel() < need:
            self._dws = torch.empty(need, dtype=torch.uint8, device=q.device)
        planes = torch.empty((6, self.spec.ny, self.spec.nx), dtype=torch.float64, device=q.device) if fields \
            else None
        out = torch.empty(_lib.DIAG_COUNT, dtype=torch.float64, device=q.device)
        check(lib().hitx_diagnostics(self.handle, _p(q), _p(mask_u8), modes.ctypes.data_as(ctypes.c_void_p),
                                     len(modes), float(drag_coef), _p(planes), _p(out), _p(self._dws),
                                     self._dws.numel(), _stream()))
        sums = out.cpu().numpy()
        return (sums, planes) if fields else sums

    def hit_initial_state(self, modes, target_mach: float, out):
        """HIT initial state into out from (kx, ky, amplitude, phase) rows;
        returns the velocity rms before scaling (synchronises)."""
        self._q(out, "out")
        modes = np.ascontiguousarray(np.asarray(modes, dtype=np.float64).reshape(-1, 4))
        need = int(lib().hitx_diag_workspace_bytes(self.handle, 0))
        ws = torch.empty(need, dtype=torch.uint8, device=out.device)  # setup only: not kept
        sc = torch.empty(2, dtype=torch.float64, device=out.device)
        check(lib().hitx_hit_initial_state(self.handle, modes.ctypes.data_as(ctypes.c_void_p), len(modes),
                                           float(target_mach), _p(out), _p(sc), _p(ws), need, _stream()))
        return float(sc[0].cpu())

    def riemann_initial_state(self, quadrants, x_min, y_min, x_split, y_split, out):
        """Quadrant initial state into out (rows (rho, u, v, p) of I, II, III, IV)."""
        self._q(out, "out")
        qd = np.ascontiguousarray(np.asarray(quadrants, dtype=np.float64).reshape(4, 4))
        check(lib().hitx_riemann_initial_state(self.handle, qd.ctypes.data_as(ctypes.c_void_p), float(x_min),
                                               float(y_min), float(x_split), float(y_split), _p(out), _stream()))
        return out

    def riemann_diagnostics(self, q, mask_u8=None, fields: bool = False):
        """HITX_RDIAG_* values of q (synchronises) and, when fields, the device
        planes [5, ny, nx] = rho, u, v, p, omega_z (float64)."""
        self._q(q)
        if mask_u8 is not None:
            require_device(mask_u8, "mask", torch.uint8, (self.spec.ny, self.spec.nx))
        planes = torch.empty((5, self.spec.ny, self.spec.nx), dtype=torch.float64, device=q.device) if fields \
            else None
        out = torch.empty(_lib.RDIAG_COUNT, dtype=torch.float64, device=q.device)
        check(lib().hitx_riemann_diagnostics(self.handle, _p(q), _p(mask_u8), _p(planes), _p(out), _p(self.ws),
                                             self.ws_bytes, _stream()))
        vals = out.cpu().numpy()
        return (vals, planes) if fields else vals

    def set_line_form(self, form: int):
        """0: one wavefront per line where the line length allows (multiple of
        64, at most 512), else the block form; 1: the block form everywhere."""
        check(lib().hitx_set_line_form(self.handle, int(form)))

    def set_sensor_form(self, form: int):
        """0: default (split); 1: fused tile; 2: raw-mask column walk + dilation tile."""
        check(lib().hitx_set_sensor_form(self.handle, int(form)))

    def timing(self, on: bool):
        check(lib().hitx_timing_enable(self.handle, int(on)))

    def timing_read(self):
        """(summed line-sweep kernel ms, launches) since the last read."""
        ms, n = ctypes.c_double(), ctypes.c_int32()
        check(lib().hitx_timing_read(self.handle, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def forcing_project(self, noise_f64):
        require_device(noise_f64, "noise", torch.float64, (self.spec.ny, self.spec.nx))
        nm = len(self.spec.force_modes) // 2
        out = torch.empty(2 * nm, dtype=torch.float64, device=noise_f64.device)
        check(lib().hitx_forcing_project(self.handle, _p(noise_f64), _p(out), _p(self.ws), self.ws_bytes,
                                         _stream()))
        return out

    def forcing_synthesize(self, fhat, fx, fy, rho, u, v):
        for a, n in ((fx, "fx"), (fy, "fy"), (rho, "rho"), (u, "u"), (v, "v")):
            self._f(a, n)
        check(lib().hitx_forcing_synthesize(self.handle, _p(fhat), _p(fx), _p(fy), _p(rho), _p(u), _p(v),
                                            _p(self.red), _p(self.ws), self.ws_bytes, _stream()))
        return self.red[:2]

    def forcing_scale(self, fx, fy, scale, alpha, rho, u, v):
        check(lib().hitx_forcing_scale(self.handle, _p(fx), _p(fy), float(scale), float(alpha), _p(rho), _p(u),
                                       _p(v), _p(self.red), _p(self.ws), self.ws_bytes, _stream()))
        return self.red[:4]


_CACHE: dict = {}


def get_plan(spec: PlanSpec) -> Plan:
    key = (spec, torch.cuda.current_device())
    p = _CACHE.get(key)
    if p is None:
        if not torch.cuda.is_available():
            raise RuntimeError("the hip backend needs a HIP device (torch.cuda.is_available() is False)")
        p = Plan(spec)
        _CACHE[key] = p
    return p
