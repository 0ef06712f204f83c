"""Toeplitz-embedded Gram operator: drop-in plan_spgram / sparse_gram.

Mirrors src/pyfourier/_sparsegram/_plan.py:20-270 and
src/pyfourier/_sparsegram/_fwd.py:17-119.

plan_spgram computes, for the non-Cartesian trajectory, the point-spread
functions on the 2N grid with the engine's own adjoint NUFFT (coordinates
2k, kernel width `width`, oversampling 2, unscaled adjoint):

    PSF_rr'(d) = sum_j w_j conj(c_r(j)) c_r'(j) exp(+2 pi i k_j . d / N),  d in [-N, N)

with c_r(j) = Bt[l, t_j] * basis[f_j, k] the per-sample subspace/B0
coefficients (1 without them), and stores T = FFT(wrap(PSF)) / Ghat, where
Ghat = prod(ceil(oversamp * N)) is the oversampled size of the NUFFT whose
normal operator the Gram replaces (`oversamp`, default 1.25 = nufft's
default), so sparse_gram(x) equals nufft_adj(w * nufft(x)) for that NUFFT up
to the kernel approximation, under either norm.  The reference instead
divides T by mean|T| (defect A.16), which makes the Gram scale arbitrary.

sparse_gram pads to 2N, FFTs, applies T (pointwise, or a per-voxel R x R
matvec for subspace/B0), inverse-FFTs and crops, all in the native engine
(pf_toeplitz_apply).  The operator is Hermitian, so its autograd backward is
itself (src/pyfourier/_sparsegram/_fwd.py:103-119).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._b0 import mri_exp_approx
from ._nufft import _Call, _to_torch
from ._plan import plan_nufft, resolve_device
from ._traj import normalize_coordinates


class NativeToeplitz:
    def __init__(self, kernel: torch.Tensor, shape, nseg: int, ncoeff: int, device):
        L = _lib.lib()
        d = len(shape)
        sh = (ctypes.c_int * d)(*[int(s) for s in shape])
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(L.pf_toeplitz_create(_lib.ptr(kernel), d, sh, int(nseg), int(ncoeff),
                                            _lib.stream_of(device), ctypes.byref(h)))
        self.handle = h
        self._lib = L

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and self._lib is not None:
            try:
                self._lib.pf_toeplitz_destroy(h)
            except Exception:
                pass
            self.handle = None


@dataclass
class GramMatrix:
    """Same fields as the reference GramMatrix (src/pyfourier/_sparsegram/_plan.py:253-270)."""

    value: object  # kernel T: (2N...) or (R, R, 2N...) complex64
    shape: tuple  # embedding grid 2N (z, y, x)
    ndim: int
    islowrank: bool
    zmap_s_kernel: object
    zmap_batch_size: int
    device: object
    image_shape: tuple = ()
    ncoeff: int = 1
    nseg: int = 1
    _native: object = field(default=None, repr=False)

    def to(self, device):
        dev = resolve_device(device)
        if self.device is None or dev != self.device:
            raise RuntimeError("a GramMatrix lives on the device it was planned on")
        return self


def plan_spgram(
    coord,
    shape,
    basis=None,
    zmap=None,
    L=6,
    nbins=(40, 40),
    dt=None,
    T=None,
    L_batch_size=None,
    weight=None,
    oversamp=None,
    width=6,
    device="cpu",
    units="auto",
):
    """Compute the Toeplitz kernel of A^H W A (same signature as the reference).

    coord (ncontrasts, nviews, nsamples, ndim); shape: image matrix; basis (ncontrasts, ncoeff);
    weight broadcastable to (ncontrasts, nviews, nsamples); device must name a GPU.
    units (extension): coordinate unit, as for plan_nufft.  The per-voxel block
    R = (field-map segments) x (subspace coefficients) is at most PF_TOEP_RMAX.
    """
    dev = resolve_device(device)
    kb = 1 if basis is None else int(_to_torch(basis).shape[-1])
    rb = kb * (int(L) if zmap is not None else 1)
    if rb > _lib.PF_TOEP_RMAX:
        raise ValueError(
            f"the Gram kernel block R = nseg * ncoeff = {rb} exceeds {_lib.PF_TOEP_RMAX}; "
            "use fewer field-map segments (L) or subspace coefficients"
        )
    coord = _to_torch(coord).to(torch.float32)
    d = int(coord.shape[-1])
    while coord.dim() < 4:
        coord = coord[None]
    F, nv, ns = (int(s) for s in coord.shape[:3])
    shp = [int(shape)] * d if np.isscalar(shape) else [int(s) for s in list(shape)[-d:]]
    kg = normalize_coordinates(coord, shp[::-1], units=units).to(dev)  # grid units
    M = F * nv * ns

    # per-sample coefficients c_r(j), r = (l, k)
    coefs = torch.ones((M, 1), dtype=torch.complex64, device=dev)
    K = 1
    if basis is not None:
        b = _to_torch(basis).to(device=dev, dtype=torch.complex64)
        if b.dim() != 2 or b.shape[0] != F:
            raise ValueError("basis must have shape (ncontrasts, ncoeff)")
        K = int(b.shape[1])
        frame = torch.arange(F, device=dev).repeat_interleave(nv * ns)
        coefs = b[frame]  # (M, K)
    nseg, zs = 1, None
    if zmap is not None:
        if T is None:
            if dt is None:
                raise ValueError("Please provide raster time dt if T is not known")
            T = dt * torch.arange(ns, dtype=torch.float32)
        bt, zs = mri_exp_approx(_to_torch(zmap), _to_torch(T), L, nbins)
        bt = bt.to(dev)
        zs = zs.to(dev)
        nseg = int(bt.shape[0])
        tidx = torch.arange(M, device=dev) % ns
        cb = bt[:, tidx].T  # (M, L)
        coefs = (cb[:, :, None] * coefs[:, None, :]).reshape(M, nseg * K)
    R = nseg * K

    w = torch.ones((M,), dtype=torch.float32, device=dev)
    if weight is not None:
        w = torch.broadcast_to(_to_torch(weight).to(device=dev, dtype=torch.float32), (F, nv, ns)).reshape(-1)

    # PSFs with the adjoint NUFFT on the 2N grid (coordinates 2k, one frame)
    big = [2 * s for s in shp]
    plan2 = plan_nufft(2.0 * kg.reshape(1, 1, M, d).cpu(), big, width=width, oversamp=2.0, units="grid")
    plan2.to(dev)
    y = w[None, :] * torch.conj(coefs.T)[:, None, :] * coefs.T[None, :, :]  # (R, R, M)
    call = _Call(plan2, dev, None, None, None)
    psf = call.adjoint(y.reshape(R * R, 1, M).contiguous(), _lib.PF_NORM_ADJOINT)  # (R*R, 1, 2N...)
    psf = psf.reshape((R, R) + tuple(big))
    ax = tuple(range(-d, 0))
    sigma = 1.25 if oversamp is None else oversamp
    sig = [float(sigma)] * d if np.isscalar(sigma) else [float(s) for s in list(sigma)[-d:]]
    ghat = float(np.prod([math.ceil(s * n) for s, n in zip(sig, shp)]))
    kern = torch.fft.fftn(torch.fft.ifftshift(psf, dim=ax), dim=ax) / ghat  # (R, R, 2N...)
    kern_native = kern.reshape(R, R, -1).permute(2, 0, 1).contiguous()  # (prod 2N, R, R)
    nat = NativeToeplitz(kern_native, shp, nseg, K, dev)
    value = kern[0, 0] if R == 1 else kern
    if L_batch_size is None and zs is not None:
        L_batch_size = nseg
    return GramMatrix(value, tuple(big), d, basis is not None, zs, L_batch_size, dev, tuple(shp), K, nseg, nat)


class _GramFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, gm):
        ctx.gm = gm
        return _apply(x, gm)

    @staticmethod
    def backward(ctx, g):
        return _apply(g.contiguous().to(torch.complex64), ctx.gm), None


def _apply(x: torch.Tensor, gm: GramMatrix) -> torch.Tensor:
    C = x.shape[0]
    out = torch.empty_like(x)
    if C == 0:
        return out
    L = _lib.lib()
    h = gm._native.handle
    n = ctypes.c_size_t()
    _lib.check(L.pf_toeplitz_workspace_size(h, int(C), ctypes.byref(n)))
    ws = torch.empty(max(1, n.value), dtype=torch.uint8, device=x.device)
    zs = None if gm.zmap_s_kernel is None else gm.zmap_s_kernel.reshape(gm.nseg, -1).contiguous()
    with torch.cuda.device(x.device):
        _lib.check(L.pf_toeplitz_apply(h, int(C), _lib.ptr(x), _lib.ptr(out), _lib.ptr(zs), _lib.ptr(ws), n.value,
                                       _lib.stream_of(x.device)))
    return out


def sparse_gram(image, gram_matrix: GramMatrix, norm=None, device="cpu", threadsperblock=128):
    """Apply the Toeplitz Gram (same signature as the reference sparse_gram).

    image (..., [ncoeff], [nz], ny, nx); `norm` does not change the operator
    (both FFT normalisations cancel); `threadsperblock` is ignored.
    """
    backend_numpy = not isinstance(image, torch.Tensor)
    x = _to_torch(image)
    idevice = x.device
    gm = gram_matrix
    dev = gm.device if device is None else resolve_device(device)
    if dev != gm.device:
        raise RuntimeError("sparse_gram must run on the device the GramMatrix was planned on")
    d = gm.ndim
    shp = tuple(gm.image_shape)
    if tuple(x.shape[-d:]) != shp:
        raise ValueError(f"image spatial shape {tuple(x.shape[-d:])} does not match {shp}")
    if gm.islowrank:
        if x.dim() < d + 1 or x.shape[-(d + 1)] != gm.ncoeff:
            raise ValueError(f"image must have an ncoeff={gm.ncoeff} axis before the spatial axes")
        batch = tuple(x.shape[: -(d + 1)])
    else:
        batch = tuple(x.shape[:-d])
    C = int(np.prod(batch)) if batch else 1
    dtype = x.dtype
    xin = x.to(device=dev, dtype=torch.complex64).reshape((C, gm.ncoeff) + shp).contiguous()
    y = _GramFn.apply(xin, gm).reshape(x.shape)
    if not dtype.is_complex:
        y = y.real
    y = y.to(dtype if (dtype.is_complex or dtype.is_floating_point) else torch.complex64)
    if backend_numpy:
        return y.detach().cpu().numpy()
    return y.to(idevice)
