"""Dense Cartesian Gram operator: drop-in plan_gram / gram.

Mirrors src/pyfourier/_gram/_plan.py:14-177 and src/pyfourier/_gram/_fwd.py:15-111
with the intent of src/pyfourier/_gram/_gram.py:11-133: the normal operator
of the masked Cartesian FFT (fftn) is

    gram(x) = sum_l' conj(Ct_l') IFFTc( S (x) FFTc(Ct_l x_k) )

with, per k-space voxel, the R x R matrix (R = L segments x K coefficients)

    S[(l',k'), (l,k)] = sum_t mask_t * conj(Bt[l', kx]) * Bt[l, kx] * basis[t, k'] * conj(basis[t, k])

(the mask alone, per contrast, without basis and field map).  Both FFTs are
orthonormal, so gram(x) = ifftn(fftn(x)) for a binary mask under either
norm, and `norm` does not change the operator.  plan_gram builds S once
(plan time, on the GPU) in uncentred FFT order; gram runs in the native
engine (pf_gram_create + pf_toeplitz_apply): the fused FFT passes on the N
grid, one kernel for the per-voxel product, the inverse passes.

Deviations (reference defects): _gram reads a field that GramMatrix does not
have (A.18) and applies the subspace matrices with a broadcast that does not
line up with the image layout; the mask is applied without the FFT-order
shift in the non-subspace case; the field map enters only as a spatial
factor.  The build follows the normal-operator definition above instead, and
accepts 1-, 2- or 3-D grids (the reference asserts 2-D).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._b0 import mri_exp_approx
from ._nufft import _to_torch
from ._plan import resolve_device


class NativeGram:
    def __init__(self, kernel: torch.Tensor, shape, nframes: int, nseg: int, ncoeff: int, device):
        L = _lib.lib()
        d = len(shape)
        sh = (ctypes.c_int * d)(*[int(s) for s in shape])
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(L.pf_gram_create(_lib.ptr(kernel), d, sh, int(nframes), int(nseg), int(ncoeff),
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
    """Same fields as the reference GramMatrix (src/pyfourier/_gram/_plan.py:163-177)."""

    value: object  # S in centred order: (nframes, *N) without basis/B0, else (nframes, *N, R, R)
    shape: tuple
    ndim: int
    islowrank: bool
    zmap_s_kernel: object
    zmap_batch_size: object
    device: object
    nframes: int = 1
    ncoeff: int = 1
    nseg: int = 1
    zmap_t_kernel: object = None  # (L, Nx) temporal interpolator the kernel was built with
    _native: object = field(default=None, repr=False)

    def to(self, device):
        dev = resolve_device(device)
        if dev != self.device:
            raise RuntimeError("a GramMatrix lives on the device it was planned on")
        return self


def plan_gram(mask, shape, basis=None, zmap=None, L=6, nbins=(40, 40), T=None, L_batch_size=None, oversamp=None,
              width=6, device="cpu"):
    """Kernel of the dense Gram operator (same signature as the reference plan_gram).

    mask: (ncontrasts, [nz], ny, nx) binary k-space mask, or any shape that
    broadcasts to it (e.g. (ncontrasts, ny, 1) for a readout-separable 3-D
    acquisition); basis (ncontrasts, ncoeff); zmap with T (readout instants
    along x) adds field-map segments.  The GPU is the mask's if it is a GPU
    tensor, else `device`.  `oversamp` and `width` are accepted and unused
    (as in the reference).
    """
    m = _to_torch(mask)
    dev = m.device if (m is not None and m.is_cuda) else resolve_device(device)
    shp = tuple(int(s) for s in ([shape] if np.isscalar(shape) else shape))
    d = len(shp)
    if m is None:
        m = torch.ones((1,) + shp)
    if torch.is_complex(m):
        raise ValueError("the k-space mask must be real")
    m = m.to(device=dev, dtype=torch.float32)
    while m.dim() < d + 1:
        m = m[None]
    m = torch.broadcast_to(m, (m.shape[0],) + shp)
    nfr = int(m.shape[0])
    ntot = int(np.prod(shp))

    bt = zs = None
    nseg = 1
    if zmap is not None:
        if T is None:
            raise ValueError("plan_gram needs the readout sampling instants T with a field map")
        T = _to_torch(T).reshape(-1)
        if T.numel() != shp[-1]:
            raise ValueError(f"T must have one entry per readout (x) point, {shp[-1]}")
        bt, zs = mri_exp_approx(_to_torch(zmap), T, L, nbins)
        bt, zs = bt.to(dev), zs.to(dev)
        nseg = int(bt.shape[0])
        if L_batch_size is None:
            L_batch_size = nseg
    b = torch.ones((1, shp[-1]), dtype=torch.complex64, device=dev) if bt is None else bt
    bb = torch.conj(b)[:, None, :] * b[None, :, :]  # (L', L, Nx)

    K = 1
    if basis is None:
        if nseg == 1:
            S = m.to(torch.complex64)[..., None, None]  # (nfr, *N, 1, 1)
        else:
            bbx = bb.permute(2, 0, 1).reshape((1,) * d + (shp[-1], nseg, nseg))  # (.., Nx, L', L)
            S = m.to(torch.complex64)[..., None, None] * bbx  # (nfr, *N, L', L)
    else:
        B = _to_torch(basis).to(device=dev, dtype=torch.complex64)
        if B.dim() != 2:
            raise ValueError("basis must have shape (ncontrasts, ncoeff)")
        Tn, K = int(B.shape[0]), int(B.shape[1])
        if nfr not in (1, Tn):
            raise ValueError(f"mask has {nfr} frames, basis {Tn} contrasts")
        mt = torch.broadcast_to(m, (Tn,) + shp).reshape(Tn, ntot).to(torch.complex64)
        BB = (B[:, :, None] * torch.conj(B)[:, None, :]).reshape(Tn, K * K)
        Sk = (mt.T @ BB).reshape(shp + (K, K))  # sum_t m_t B[t,k'] conj(B[t,k])
        bbx = bb.permute(2, 0, 1).reshape((1,) * (d - 1) + (shp[-1], nseg, 1, nseg, 1))
        S = (bbx * Sk.reshape(shp + (1, K, 1, K))).reshape((1,) + shp + (nseg * K, nseg * K))
        nfr = 1
    R = nseg * K
    if R > 16:
        raise ValueError("the dense Gram supports nseg * ncoeff <= 16")
    S = S.contiguous()
    ax = tuple(range(1, d + 1))
    kern = torch.fft.ifftshift(S, dim=ax).reshape(nfr, ntot, R, R).contiguous()
    nat = NativeGram(kern, shp, nfr, nseg, K, dev)
    value = S[..., 0, 0] if R == 1 else S
    return GramMatrix(value, shp, d, basis is not None, zs, L_batch_size, dev, nfr, K, nseg, bt, nat)


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


class _GramFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, gm):
        ctx.gm = gm
        return _apply(x, gm)

    @staticmethod
    def backward(ctx, g):
        return _apply(g.contiguous().to(torch.complex64), ctx.gm), None


def gram(image, gram_matrix: GramMatrix, norm=None, device="cpu"):
    """Apply the dense Gram operator (same signature as the reference gram).

    image (..., [ncontrasts | ncoeff], [nz], ny, nx); `norm` does not change
    the operator (both FFT normalisations cancel).
    """
    backend_numpy = not isinstance(image, torch.Tensor)
    x = _to_torch(image)
    idevice = x.device
    gm = gram_matrix
    dev = gm.device if device is None else resolve_device(device)
    if dev != gm.device:
        raise RuntimeError("gram must run on the device the GramMatrix was planned on")
    d = gm.ndim
    shp = tuple(gm.shape)
    if tuple(x.shape[-d:]) != shp:
        raise ValueError(f"image spatial shape {tuple(x.shape[-d:])} does not match {shp}")
    planes = gm.ncoeff if gm.islowrank else gm.nframes
    if gm.islowrank or gm.nframes > 1:
        if x.dim() < d + 1 or x.shape[-(d + 1)] != planes:
            raise ValueError(f"image must have {planes} planes before the spatial axes")
        batch = tuple(x.shape[: -(d + 1)])
    else:
        batch = tuple(x.shape[:-d])
    C = int(np.prod(batch)) if batch else 1
    dtype = x.dtype
    xin = x.to(device=dev, dtype=torch.complex64).reshape((C, planes) + shp).contiguous()
    y = _GramFn.apply(xin, gm).reshape(x.shape)
    if not dtype.is_complex:
        y = y.real
    y = y.to(dtype if (dtype.is_complex or dtype.is_floating_point) else torch.complex64)
    if backend_numpy:
        return y.detach().cpu().numpy()
    return y.to(idevice)
