"""Sparse FFT: drop-in plan_spfft / sparse_fftn / sparse_ifftn.

Mirrors src/pyfourier/_sparsefft/_plan.py:15-146, _fwd.py:18-237,
_adj.py:18-234 and the operator of _sparsefft.py:11-144: a centred FFT
followed by picking the sampled Cartesian points (the dense-to-sparse masks of
src/pyfourier/_subroutines/_mask/_dense2sparse), and zero-filling plus the
inverse FFT (_sparse2dense) in the other direction, with the same subspace,
B0 and density-weight options as the NUFFT.

That operator is the NUFFT with a one-tap kernel on every axis (width 1,
oversampling 1: the tap weight is 1 and the scale 1), so a sparse plan is a
native NUFFT plan built from the integer indexes and runs through the same
kernels: the pruned fused FFT passes, the tile-sorted gather (pick) and the
deterministic spread (zero-fill with repeated points summed).  The subspace
convention is the reference's (_dense2sparse uses basis[frame, coeff],
_sparse2dense its conjugate), which is the NUFFT's.

Deviations: index normalisation follows the intent of
normalize_coordinates(indexes, shape, center=False) (defect A.11 fixed:
integer indexes in [0, N) are used as they are, centred indexes in
[-N/2, N/2) are shifted by N // 2); indexes outside the matrix wrap
periodically instead of reading out of bounds; B0 segments are added once
(A.13); autograd backward passes are exact adjoints (A.22).
"""

from __future__ import annotations

import numpy as np
import torch

from ._nufft import _to_torch, nufft, nufft_adj
from ._plan import plan_nufft


def plan_spfft(indexes, shape, zmap=None, L=6, nbins=(40, 40), dt=None, T=None, L_batch_size=None):
    """Precompute a sparse FFT plan (same signature as the reference plan_spfft).

    indexes: sampled Cartesian points (ncontrasts, nviews, nsamples, ndim)
    (or fewer leading axes), last axis (x, y[, z]), integers in [0, N) or
    centred in [-N/2, N/2); shape: matrix (scalar or (z, y, x)).  Returns a
    NUFFTPlan (usable as `plan` by sparse_fftn / sparse_ifftn) with the
    reference FFTPlan's `index`, `dshape` and `ishape` attributes.
    """
    idx = _to_torch(indexes)
    if torch.is_complex(idx):
        raise TypeError("indexes must be real")
    ndim = int(idx.shape[-1])
    shp = [int(shape)] * ndim if np.isscalar(shape) else [int(s) for s in list(shape)[-ndim:]]
    c = idx.detach().cpu().to(torch.float64)
    nxyz = torch.tensor(shp[::-1], dtype=torch.float64)
    if not bool((c < 0).any()):
        c = c - torch.floor(nxyz / 2)
    if not bool(torch.allclose(c, torch.round(c))):
        raise ValueError(
            "Input coordinates must lie on Cartesian grid, got non-uniform coord! Please use NUFFT instead."
        )
    plan = plan_nufft(torch.round(c).to(torch.float32), shp, width=1, oversamp=1.0, zmap=zmap, L=L, nbins=nbins,
                      dt=dt, T=T, L_batch_size=L_batch_size, units="grid")
    it = plan.interpolator
    cidx = (torch.round(c) + torch.floor(nxyz / 2)).to(torch.int64)
    cidx = cidx.reshape((it.nframes,) + tuple(it.dshape) + (ndim,))
    plan.index = tuple(cidx[..., a].reshape(it.nframes, -1) for a in range(ndim))[::-1]  # (z, y, x)
    plan.dshape = tuple(it.dshape)
    plan.ishape = tuple(shp)
    return plan


def sparse_fftn(image, indexes=None, shape=None, plan=None, basis=None, zmap=None, L=6, nbins=(40, 40), dt=None,
                T=None, L_batch_size=None, weight=None, norm=None, device="cpu", threadsperblock=128):
    """N-dimensional sparse FFT (drop-in for pyfourier.sparse_fftn).

    image (..., [ncontrasts | ncoeff], [nz], ny, nx) -> kspace (..., [ncontrasts], nviews, nsamples).
    """
    if plan is None:
        if indexes is None:
            raise ValueError("provide indexes or plan")
        idx = _to_torch(indexes)
        if shape is None:
            shape = tuple(_to_torch(image).shape[-int(idx.shape[-1]):])
        plan = plan_spfft(idx, shape, zmap, L, nbins, dt, T, L_batch_size)
    return nufft(image, nufft_plan=plan, basis=basis, weight=weight, norm=norm, device=device)


def sparse_ifftn(kspace, indexes=None, shape=None, plan=None, basis=None, zmap=None, L=6, nbins=(40, 40), dt=None,
                 T=None, L_batch_size=None, weight=None, norm=None, device="cpu", threadsperblock=128):
    """N-dimensional inverse sparse FFT (drop-in for pyfourier.sparse_ifftn).

    kspace (..., [ncontrasts], nviews, nsamples) -> image (..., [ncontrasts | ncoeff], [nz], ny, nx).
    """
    if plan is None:
        if indexes is None or shape is None:
            raise ValueError("provide indexes and shape, or plan")
        plan = plan_spfft(indexes, shape, zmap, L, nbins, dt, T, L_batch_size)
    return nufft_adj(kspace, nufft_plan=plan, basis=basis, weight=weight, norm=norm, device=device)
