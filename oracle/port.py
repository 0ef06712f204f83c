"""Reference NUFFT pipeline on the CPU: Numba loops restated in C/OpenMP
(oracle/c/restate.c) plus the reference's own torch CPU ops for
apodisation, zero-pad/crop and the centered FFT.

TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py and a parity checker
for tests.  Follows src/pyfourier/_nufft/_nufft.py:218-287 step by step, in
complex64 like the reference (float32 tables from oracle.restate.Plan).
"""

from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

from . import restate as R

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librestate.so")
        if not os.path.exists(path):
            from .build_port import build

            build(verbose=False)
        h = ctypes.CDLL(path)
        args = [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        args += [ctypes.c_void_p] * 6 + [ctypes.c_int] * 3 + [ctypes.c_void_p]
        for fn in (h.rs_gather, h.rs_spread):
            fn.argtypes = args
            fn.restype = None
        h.rs_set_threads.argtypes = [ctypes.c_int]
        _LIB = h
    return _LIB


def set_threads(n: int) -> None:
    lib().rs_set_threads(int(n))
    torch.set_num_threads(int(n))


class PortPlan:
    """Tables of the reference Interpolator (z, y, x order, unit axes padded)."""

    def __init__(self, plan: R.Plan):
        self.plan = plan
        d = plan.ndim
        idx = plan.index[::-1]  # (z, y, x)
        val = plan.value[::-1]
        pad = 3 - d
        one_i = np.zeros((plan.nframes * plan.npts, 1), dtype=np.int32)
        one_v = np.ones((plan.nframes * plan.npts, 1), dtype=np.float32)
        self.idx = [one_i] * pad + [np.ascontiguousarray(i, dtype=np.int32) for i in idx]
        self.val = [one_v] * pad + [np.ascontiguousarray(v, dtype=np.float32) for v in val]
        self.gshape = [1] * pad + list(plan.os_shape)
        self.widths = [a.shape[1] for a in self.idx]

    def _call(self, fn, nframes, batch, src, dst):
        p = self.plan
        fn(nframes, batch, p.npts, src.ctypes.data, *self.gshape,
           *[a.ctypes.data for a in self.idx], *[a.ctypes.data for a in self.val], *self.widths, dst.ctypes.data)


def _apodize(x: torch.Tensor, plan: R.Plan) -> torch.Tensor:
    # src/pyfourier/_subroutines/_apod.py:24-46 in float32 torch ops (pure)
    out = x
    for n in range(1, plan.ndim + 1):
        ax = -n
        if plan.width[ax] == 1:
            continue
        i = out.shape[ax]
        os_i = int(np.ceil(np.float32(plan.oversamp[ax]) * np.float32(i)))  # float32 product, as _apod.py:27
        idx = torch.arange(i, dtype=torch.float32)
        a = (plan.beta[ax] ** 2 - (math.pi * plan.width[ax] * (idx - i // 2) / os_i) ** 2) ** 0.5
        a = a / torch.sinh(a)
        a = a / a[i // 2]
        a = torch.nan_to_num(a, nan=1.0)
        out = out * a.reshape([i] + [1] * (n - 1))
    return out


def _resize(x: torch.Tensor, oshape) -> torch.Tensor:
    d = len(oshape)
    out = torch.zeros(x.shape[:-d] + tuple(oshape), dtype=x.dtype)
    isl, osl = [], []
    for i, o in zip(x.shape[-d:], oshape):
        si, so = max(i // 2 - o // 2, 0), max(o // 2 - i // 2, 0)
        c = min(i - si, o - so)
        isl.append(slice(si, si + c))
        osl.append(slice(so, so + c))
    out[(Ellipsis,) + tuple(osl)] = x[(Ellipsis,) + tuple(isl)]
    return out


def _fftc(x: torch.Tensor, d: int, norm, inverse=False) -> torch.Tensor:
    ax = tuple(range(-d, 0))
    f = torch.fft.ifftn if inverse else torch.fft.fftn
    return torch.fft.fftshift(f(torch.fft.ifftshift(x, dim=ax), dim=ax, norm=norm), dim=ax)


def nufft(x: np.ndarray, pp: PortPlan, norm=None) -> np.ndarray:
    """x (B, F, *N) complex64 -> (B, F, M) complex64 (no basis, no field map)."""
    p = pp.plan
    B, F = x.shape[:2]
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex64))
    t = _apodize(t, p)
    t = _resize(t, p.os_shape)
    t = _fftc(t, p.ndim, norm)
    g = np.ascontiguousarray(t.transpose(0, 1).numpy())  # (F, B, G...) transpose-copy
    out = np.zeros((F, B, p.npts), dtype=np.complex64)
    pp._call(lib().rs_gather, F, B, g, out)
    return (out / np.float32(p.scale)).transpose(1, 0, 2)


def nufft_adj(y: np.ndarray, pp: PortPlan, norm=None) -> np.ndarray:
    """y (B, F, M) complex64 -> (B, F, *N) complex64."""
    p = pp.plan
    B, F = y.shape[:2]
    k = np.ascontiguousarray(np.asarray(y, dtype=np.complex64).transpose(1, 0, 2))
    g = np.zeros((F, B) + tuple(p.os_shape), dtype=np.complex64)
    pp._call(lib().rs_spread, F, B, k, g)
    t = torch.from_numpy(g).transpose(0, 1) / p.scale
    t = _fftc(t, p.ndim, norm, inverse=True)
    t = _resize(t, p.shape)
    return _apodize(t, p).numpy()
