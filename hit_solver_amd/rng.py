"""Forcing noise on the device: NumPy Generator(PCG64).standard_normal,
reproduced bit for bit by ``hitx_normals`` (the reference draws it on the host,
OOP/forcing.py:66 and :77)."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib

_MASK64 = (1 << 64) - 1


def pcg64_words(bit_generator) -> np.ndarray:
    """PCG64 state and increment of a NumPy bit generator as 4 uint64 words
    (state high, state low, increment high, increment low)."""
    st = bit_generator.state
    if st.get("bit_generator") != "PCG64":
        raise ValueError("the device stream reproduces PCG64 only")
    s, inc = int(st["state"]["state"]), int(st["state"]["inc"])
    return np.array([s >> 64, s & _MASK64, inc >> 64, inc & _MASK64], dtype=np.uint64)


_MASK128 = (1 << 128) - 1
# the PCG64 LCG multiplier (O'Neill's 128-bit default, as NumPy's PCG64 uses)
_PCG_MULT = 0x2360ED051FC65DA44385DF649FCCF645


def pcg64_advance(state: int, inc: int, delta: int) -> int:
    """State of the PCG64 LCG s -> s*M + inc after `delta` steps (delta taken
    mod 2^128, so a negative delta steps back), by the square-and-multiply
    jump of O'Neill's PCG paper (Brown's arbitrary-stride LCG skip)."""
    delta &= _MASK128
    cm, cp, am, ap = _PCG_MULT, inc & _MASK128, 1, 0
    while delta:
        if delta & 1:
            am = (am * cm) & _MASK128
            ap = (ap * cm + cp) & _MASK128
        cp = ((cm + 1) * cp) & _MASK128
        cm = (cm * cm) & _MASK128
        delta >>= 1
    return (am * state + ap) & _MASK128


def pcg64_words_rewind(words, k: int) -> np.ndarray:
    """4-word PCG64 state (state high, state low, increment high, increment
    low) moved back by k 64-bit outputs."""
    w = np.asarray(words).view(np.uint64)
    s, inc = (int(w[0]) << 64) | int(w[1]), (int(w[2]) << 64) | int(w[3])
    s = pcg64_advance(s, inc, -int(k))
    return np.array([s >> 64, s & _MASK64, w[2], w[3]], dtype=np.uint64)


class DeviceNormalStream:
    """The stream of ``numpy.random.default_rng(seed).standard_normal`` with
    its state on the device.  ``normals(shape)`` returns the next values as a
    float64 CUDA tensor and advances the device state; no host round trip."""

    def __init__(self, seed=None, *, generator=None, device="cuda"):
        g = generator if generator is not None else np.random.default_rng(seed)
        words = pcg64_words(g.bit_generator)
        self.state = torch.from_numpy(words.view(np.int64).copy()).to(device)
        self._ws = None
        self._ws_n = 0

    def _workspace(self, n):
        if self._ws is None or self._ws_n != n:
            nbytes = int(lib().hitx_normals_workspace_bytes(n))
            if nbytes == 0:
                raise ValueError("normal count out of range (1 .. 2^30)")
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.state.device)
            self._ws_n = n
        return self._ws

    def normals(self, shape, out=None, serial=False):
        shape = (shape,) if isinstance(shape, int) else tuple(shape)
        n = int(np.prod(shape))
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=self.state.device)
        elif out.dtype != torch.float64 or not out.is_contiguous() or out.numel() != n:
            raise ValueError("out must be a contiguous float64 tensor with the requested size")
        ws = self._workspace(n)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(lib().hitx_normals(ctypes.c_void_p(self.state.data_ptr()), ctypes.c_void_p(out.data_ptr()), n,
                                 1 if serial else 0, ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream))
        return out

    def host_state(self):
        """(state, inc) as Python ints (synchronises)."""
        w = self.state.cpu().numpy().view(np.uint64)
        return (int(w[0]) << 64) | int(w[1]), (int(w[2]) << 64) | int(w[3])
