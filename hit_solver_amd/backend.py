"""Backend selection: the new ``hip`` backend beside the reference's numpy/cupy.

Mirrors reference OOP/parallel/backend.py:8-39 (``array_module``,
``select_array_module``, ``to_numpy``).  This package implements the ``hip``
backend only; ``numpy`` and ``cupy`` remain the reference's own code paths.
"""

from __future__ import annotations

import numpy as np


class _HipModule:
    """Marker returned by :func:`select_array_module` for the hip backend."""

    __name__ = "hip"

    def __repr__(self):
        return "<array backend hip (MI355X HIP kernels, torch device buffers)>"


HIP = _HipModule()


def hip_available() -> bool:
    try:
        import torch

        from ._lib import LIB_PATH

        return LIB_PATH.exists() and torch.cuda.is_available()
    except Exception:
        return False


def select_array_module(backend: str):
    """OOP/parallel/backend.py:17-32 plus "hip".  "auto" prefers hip."""
    if backend == "hip":
        from ._lib import lib

        lib()  # RuntimeError when the library is not built
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("HIP backend requested, but no HIP device is visible.")
        return HIP
    if backend == "auto":
        return HIP if hip_available() else np
    if backend == "numpy":
        return np
    if backend == "cupy":
        raise RuntimeError("CuPy backend requested, but cupy is not installed.")
    raise ValueError("backend must be 'auto', 'numpy', 'cupy' or 'hip'")


def array_module(values):
    """OOP/parallel/backend.py:8-14: the backend an array belongs to."""
    mod = type(values).__module__.split(".")[0]
    if mod == "torch" and getattr(values, "is_cuda", False):
        return HIP
    return np


def to_numpy(values):
    """OOP/parallel/backend.py:35-39."""
    if array_module(values) is HIP:
        return values.detach().cpu().numpy()
    return values
