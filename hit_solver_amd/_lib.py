"""ctypes binding of libhitx.so (the C ABI declared in include/hitx.h).

The library is built in-tree (``hit_solver_amd/libhitx.so``) by
``hit_solver_amd.build.build()``.  There is no fallback: if the library is
missing, :func:`lib` raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhitx.so"

HITX_ABI_VERSION = 1
HITX_F64, HITX_F32 = 0, 1
HITX_HYBRID, HITX_WENO = 0, 1
HITX_EINVAL, HITX_EDEVICE, HITX_ENOMEM = -1, -2, -3
HITX_MAX_MODES = 128
# hitx_diagnostics output slots (include/hitx.h HITX_DIAG_*)
(DIAG_RHO, DIAG_U, DIAG_V, DIAG_P, DIAG_C, DIAG_U2V2, DIAG_RUU, DIAG_RVV, DIAG_RUV, DIAG_RKE, DIAG_DIV2,
 DIAG_VORT2, DIAG_LOW, DIAG_DRAG_E, DIAG_MASK, DIAG_COUNT) = range(16)
# hitx_riemann_diagnostics output slots (HITX_RDIAG_*)
(RDIAG_RHO_MIN, RDIAG_RHO_MAX, RDIAG_P_MIN, RDIAG_P_MAX, RDIAG_VORT_MAX, RDIAG_NONFINITE_Q, RDIAG_NONFINITE_RP,
 RDIAG_RHO_NONPOS, RDIAG_P_NONPOS, RDIAG_MASK, RDIAG_COUNT) = range(11)

c_i32 = ctypes.c_int32
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p


class HitxDesc(ctypes.Structure):
    """Mirror of ``hitx_desc`` (include/hitx.h)."""

    _fields_ = [
        ("abi_version", c_i32), ("nx", c_i32), ("ny", c_i32), ("periodic", c_i32),
        ("dtype", c_i32), ("scheme", c_i32),
        ("dx", c_dbl), ("dy", c_dbl), ("dx2", c_dbl), ("dy2", c_dbl),
        ("gamma", c_dbl), ("rho_floor", c_dbl), ("p_floor", c_dbl),
        ("sensor_width", c_i32), ("boundary_guard", c_i32), ("shear_enabled", c_i32), ("pad0", c_i32),
        ("compression_threshold", c_dbl), ("jump_threshold", c_dbl), ("shear_threshold", c_dbl),
        ("viscosity", c_dbl), ("prandtl", c_dbl), ("gas_constant", c_dbl),
        ("hv_coef", c_dbl * 4),
        ("guard_cells", c_i32), ("drag_enabled", c_i32), ("cooling_enabled", c_i32),
        ("forcing_enabled", c_i32),
        ("drag_coef", c_dbl), ("cooling_time", c_dbl), ("cooling_p_ref", c_dbl),
        ("n_drag_modes", c_i32), ("n_force_modes", c_i32),
        ("drag_modes_host", ctypes.POINTER(c_i32)), ("force_modes_host", ctypes.POINTER(c_i32)),
    ]


# name -> (restype, argtypes); every symbol include/hitx.h declares
SIGNATURES = {
    "hitx_plan_create": (c_i32, [ctypes.POINTER(HitxDesc), ctypes.POINTER(c_vp)]),
    "hitx_plan_destroy": (c_i32, [c_vp]),
    "hitx_workspace_bytes": (ctypes.c_size_t, [c_vp]),
    "hitx_last_error": (ctypes.c_char_p, []),
    "hitx_abi_version": (c_i32, []),
    "hitx_sensor": (c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_rhs": (c_i32, [c_vp, c_vp, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_rk3_stage": (c_i32, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_rk3_stage_static": (c_i32, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp]),
    "hitx_hyperviscosity": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_vp, c_i32, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_enforce": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_vp]),
    "hitx_step_reduce": (c_i32, [c_vp, c_vp, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_forcing_project": (c_i32, [c_vp, c_vp, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_forcing_synthesize": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                        ctypes.c_size_t, c_vp]),
    "hitx_forcing_scale": (c_i32, [c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp,
                                   ctypes.c_size_t, c_vp]),
    "hitx_primitives": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "hitx_viscous_rhs": (c_i32, [c_vp, c_vp, c_vp, c_vp]),
    "hitx_diag_workspace_bytes": (ctypes.c_size_t, [c_vp, c_i32]),
    "hitx_diagnostics": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i32, c_dbl, c_vp, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_riemann_diagnostics": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_hit_initial_state": (c_i32, [c_vp, c_vp, c_i32, c_dbl, c_vp, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_riemann_initial_state": (c_i32, [c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_vp]),
    "hitx_timing_enable": (c_i32, [c_vp, c_i32]),
    "hitx_set_line_form": (c_i32, [c_vp, c_i32]),
    "hitx_set_sensor_form": (c_i32, [c_vp, c_i32]),
    "hitx_normals_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int64]),
    "hitx_normals": (c_i32, [c_vp, c_vp, ctypes.c_int64, c_i32, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_post_workspace_bytes": (ctypes.c_size_t, [c_i32, c_i32, c_i32]),
    "hitx_post_fields": (c_i32, [c_i32, c_i32, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp,
                                 c_vp, ctypes.c_size_t, c_vp]),
    "hitx_dft2": (c_i32, [c_i32, c_i32, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp, ctypes.c_size_t, c_vp]),
    "hitx_shell_spectra": (c_i32, [c_i32, c_i32, c_dbl, c_dbl, c_vp, c_i32, c_vp, c_vp, c_vp, ctypes.c_size_t,
                                   c_vp]),
    "hitx_correlations": (c_i32, [c_i32, c_i32, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, ctypes.c_size_t,
                                  c_vp]),
    "hitx_batch_create": (c_i32, [ctypes.POINTER(HitxDesc), c_vp, c_i32, ctypes.POINTER(c_vp)]),
    "hitx_batch_destroy": (c_i32, [c_vp]),
    "hitx_batch_workspace_bytes": (ctypes.c_size_t, [c_vp]),
    "hitx_batch_bind": (c_i32, [c_vp, c_vp, ctypes.c_size_t]),
    "hitx_batch_buffer": (c_i32, [c_vp, c_i32, c_i32, ctypes.POINTER(c_vp)]),
    "hitx_batch_fixed_dt": (c_i32, [c_vp, c_vp]),
    "hitx_batch_tune": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_i32]),
    "hitx_batch_stage_form": (c_i32, [c_vp, c_i32]),
    "hitx_batch_forms": (c_i32, [c_vp, c_vp]),
    "hitx_batch_serial_draws": (c_i32, [c_vp, c_i32]),
    "hitx_batch_timing_enable": (c_i32, [c_vp, c_i32]),
    "hitx_batch_timing_read": (c_i32, [c_vp, ctypes.POINTER(c_dbl), ctypes.POINTER(c_i32)]),
    "hitx_batch_time_updates": (c_i32, [c_vp, c_i32, c_vp, ctypes.POINTER(c_dbl)]),
    "hitx_batch_step": (c_i32, [c_vp, c_vp]),
    "hitx_timing_read": (c_i32, [c_vp, ctypes.POINTER(c_dbl), ctypes.POINTER(c_i32)]),
}

_LIB = None


def lib():
    """Load libhitx.so once; raise RuntimeError when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = Path(os.environ.get("HITX_LIB", LIB_PATH))
    if not path.exists():
        raise RuntimeError(
            f"HIP backend library {path} is not built; run `python -c \"import __graft_entry__ as g; g.build()\"`"
        )
    h = ctypes.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    if h.hitx_abi_version() != HITX_ABI_VERSION:
        raise RuntimeError("libhitx.so ABI version mismatch; rebuild the library")
    _LIB = h
    return h


def check(rc: int) -> None:
    """Map a hitx return code to the reference's exception types."""
    if rc == 0:
        return
    msg = (lib().hitx_last_error() or b"").decode(errors="replace")
    if rc in (HITX_EINVAL, HITX_ENOMEM):
        raise ValueError(msg)
    raise RuntimeError(msg)
