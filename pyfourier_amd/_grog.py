"""GRAPPA operator gridding: drop-in grog_interp.

Mirrors src/pyfourier/_preinterp/_grog.py:22-249.  The small per-call linear
algebra (operator training by regularised least squares, fractional operator
powers through an eigendecomposition) runs on the host in complex128; the
per-sample work runs in the native engine (pf_grog_index, pf_grog_apply):
quantised shifts and grid indexes, then every sample's coil vector moved onto
its nearest Cartesian point by the per-axis operator powers.

Deviations (reference defects, the intent is implemented):

* the training and power helpers take their array module from the builtin
  `input` (:276, :311, :341, :347), so they raise; here they use the
  calibration data;
* _dot_product (:406-413) multiplies out[j] += G[i][j] * x[j] (the input's
  own coil j, a per-column scaling); the operator was trained as T = S @ G,
  so the intended map is the row-vector product out[j] = sum_i x[i] G[i][j];
* the 3-D path reshapes the samples to (1, npts, nc) and then reads one shift
  index for all of them (:198-200 with :415-423); here every sample uses its
  own shifts, and a 2-D operator set broadcasts over batch entries when the
  calibration has a single slice;
* the shift quantisation multiplies by (nsteps - 1) where the reference
  hard-codes 10 (:178) -- identical for the default nsteps = 11, and in range
  for other odd nsteps.

Returns (output, indexes, weights) like the reference: the gridded samples
(zero where the nearest grid point is outside the matrix), the grid index per
sample (x, y[, z], clamped into the matrix) and 1 / (number of samples that
share the sample's grid point).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._nufft import _to_torch
from ._plan import resolve_device


def _lstsq_op(S: np.ndarray, T: np.ndarray, lamda: float) -> np.ndarray:
    """G = (S^H S + lamda0 I)^-1 S^H T with lamda0 = lamda ||S^H||_F / nc
    (src/pyfourier/_preinterp/_grog.py:291-305,330-340)."""
    Sh = np.conj(S).swapaxes(-1, -2)
    nc = S.shape[-1]
    lam0 = lamda * np.linalg.norm(Sh.reshape(Sh.shape[:-2] + (-1,)), axis=-1) / nc
    A = Sh @ S + lam0[..., None, None] * np.eye(nc)
    return np.linalg.solve(A, Sh @ T)


def train_operators(calib: np.ndarray, ndim: int, lamda: float) -> list:
    """GRAPPA shift operators per axis, ordered (x, y[, z]).

    calib (nc, [nz], ny, nx); 2-D: one operator set per calibration slice,
    each (nslices, nc, nc); 3-D: (1, nc, nc)."""
    c = np.asarray(calib, dtype=np.complex128)
    if c.ndim == 3:
        c = c[:, None]
    c = np.moveaxis(c, 0, -1)  # (nz, ny, nx, nc)
    nz, ny, nx, nc = c.shape
    if ndim == 2:
        gy = _lstsq_op(c[:, :-1].reshape(nz, -1, nc), c[:, 1:].reshape(nz, -1, nc), lamda)
        gx = _lstsq_op(c[:, :, :-1].reshape(nz, -1, nc), c[:, :, 1:].reshape(nz, -1, nc), lamda)
        return [gx, gy]
    if ndim != 3:
        raise ValueError("GROG needs 2-D or 3-D coordinates")
    gz = _lstsq_op(c[:-1].reshape(1, -1, nc), c[1:].reshape(1, -1, nc), lamda)
    gy = _lstsq_op(c[:, :-1].reshape(1, -1, nc), c[:, 1:].reshape(1, -1, nc), lamda)
    gx = _lstsq_op(c[:, :, :-1].reshape(1, -1, nc), c[:, :, 1:].reshape(1, -1, nc), lamda)
    return [gx, gy, gz]


def operator_powers(G: np.ndarray, deltas: np.ndarray) -> np.ndarray:
    """G^delta = V diag(L^delta) V^-1 (principal powers), (nsteps, ...) for
    G (..., nc, nc) (src/pyfourier/_preinterp/_grog.py:344-367)."""
    L, V = np.linalg.eig(G)
    Vi = np.linalg.inv(V)
    P = L[None] ** np.asarray(deltas, dtype=np.float64).reshape((-1,) + (1,) * L.ndim)
    return (V[None] * P[..., None, :]) @ Vi[None]


def grog_interp(input, calib, coord, shape, lamda=0.01, nsteps=11, device=None, threadsperblock=128):
    """GROG interpolation of non-Cartesian k-space (same signature as the reference).

    input (..., nc, ncontrasts, nviews, nsamples) (trajectory axes as in
    coord); calib (nc, [nz], ny, nx); coord (ncontrasts, nviews, nsamples,
    ndim) in grid units (-N/2, N/2); shape: matrix (scalar or (z, y, x)).
    Runs on `device` (default: the input's GPU, or the current GPU for host
    inputs); `threadsperblock` is accepted and ignored.
    """
    backend_numpy = not isinstance(input, torch.Tensor)
    x = _to_torch(input)
    idevice = x.device
    if device is None:
        dev = resolve_device(idevice if x.is_cuda else torch.device("cuda", torch.cuda.current_device()))
    else:
        dev = resolve_device(device)
    c = _to_torch(coord).detach().to(torch.float32)
    cshape = tuple(c.shape)
    ndim = int(cshape[-1])
    if ndim not in (2, 3):
        raise ValueError("GROG needs 2-D or 3-D coordinates")
    npts = int(np.prod(cshape[:-1]))
    shp = [int(shape)] * ndim if np.isscalar(shape) else [int(s) for s in list(shape)[-ndim:]]
    ntraj = len(cshape) - 1
    if x.dim() < ntraj + 1 or tuple(x.shape[-ntraj:]) != cshape[:-1]:
        raise ValueError(f"input {tuple(x.shape)} does not end with the trajectory shape {cshape[:-1]}")
    nc = int(x.shape[-ntraj - 1])
    B = int(np.prod(x.shape[: -ntraj - 1])) if x.dim() > ntraj + 1 else 1

    cal = np.asarray(_to_torch(calib).detach().cpu().to(torch.complex128).numpy())
    if cal.shape[0] != nc:
        raise ValueError(f"calibration has {cal.shape[0]} coils, the data {nc}")
    cal = cal / np.linalg.norm(cal)
    nsteps = 2 * (int(nsteps) // 2) + 1
    ops = train_operators(cal, ndim, lamda)
    deltas = (np.arange(nsteps) - (nsteps - 1) // 2) / (nsteps - 1)
    tables = np.stack([operator_powers(g, deltas) for g in ops])  # (ndim, nsteps, nob, nc, nc)
    nob = int(tables.shape[2])
    if nob not in (1, B):
        raise ValueError(f"{nob} calibration slices for {B} batch entries")
    tab = torch.from_numpy(tables.astype(np.complex64)).to(dev).contiguous()

    L = _lib.lib()
    cd = c.to(dev).reshape(npts, ndim).contiguous()
    lut = torch.empty(npts, dtype=torch.int32, device=dev)
    index = torch.empty((npts, ndim), dtype=torch.int32, device=dev)
    flat = torch.empty(npts, dtype=torch.int64, device=dev)
    xin = x.to(device=dev, dtype=torch.complex64).reshape(B, nc, npts).contiguous()
    out = torch.empty_like(xin)
    sh = (ctypes.c_int * ndim)(*shp)
    stream = _lib.stream_of(dev)
    with torch.cuda.device(dev):
        _lib.check(L.pf_grog_index(_lib.ptr(cd), ndim, sh, npts, nsteps, _lib.ptr(lut), _lib.ptr(index),
                                   _lib.ptr(flat), stream))
        _lib.check(L.pf_grog_apply(_lib.ptr(xin), _lib.ptr(out), B, nc, npts, ndim, nsteps, _lib.ptr(tab), nob,
                                   _lib.ptr(lut), stream))
    _, inv, counts = torch.unique(flat, return_inverse=True, return_counts=True)
    weights = (1.0 / counts[inv].to(torch.float32)).reshape(cshape[:-1])
    output = out.reshape(x.shape)
    indexes = index.to(torch.int64).reshape(cshape)
    if backend_numpy:
        return output.cpu().numpy(), indexes.cpu().numpy(), weights.cpu().numpy()
    return output.to(idevice), indexes.to(idevice), weights.to(idevice)
