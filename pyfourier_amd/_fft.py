"""Cartesian (masked) FFT: drop-in plan_fft / fftn / ifftn and fft / ifft.

Mirrors src/pyfourier/_fft/_plan.py:14-108 (plan_fft), _fwd.py:18-196 (fftn),
_adj.py:18-199 (ifftn), the operator of _fft.py:10-151 and the centred FFT
utilities of src/pyfourier/_subroutines/_fft.py:11-194.  Every call runs in
the native engine (pf_cart_fft / pf_cart_ifft, include/pfhip.h): the fused
FFT passes of the NUFFT on an oversampling-1 grid, then one streaming kernel
that applies the centring shift, the subspace basis, the B0 segment sum and
the mask while it moves coils out of the grid's innermost position.

Operator (centred order, image (..., [K | T], *N) -> k-space (..., [T], *N)):

    fftn   y_t = mask_t * sum_l Bt[l, kx] * sum_k conj(basis[t, k]) * FFTc(Ct_l * X_k)
    ifftn  X_k = sum_l conj(Ct_l) * IFFTc(sum_t basis[t, k] * conj(Bt[l, kx]) * mask_t * y_t)

with the B0 temporal interpolator Bt running along the readout (last) axis;
the reference's conjugate-basis forward (_basis_adj, _fft.py:141-151) is kept.
Deviations (reference defects): each B0 segment is added once (A.13); the
autograd backward is the exact adjoint of the direction in use, also under
norm=None (A.22 for this operator); a dense mask is handled at all (the
dense branch of FFTPlan, _subroutines/_mask/_plan.py:51-52, names an
undefined variable).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._b0 import mri_exp_approx
from ._nufft import _adjoint_code, _norm_code, _restore, _to_torch
from ._plan import resolve_device

_RMAX = 32  # nseg * ncoeff per native call (pf_cart_fft)


class NativeCart:
    """Owns one pf_cart (geometry + mask) on one device."""

    def __init__(self, mask: torch.Tensor | None, shape, device):
        L = _lib.lib()
        d = len(shape)
        sh = (ctypes.c_int * d)(*[int(s) for s in shape])
        h = ctypes.c_void_p()
        mframes = 0
        if mask is not None:
            mask = mask.to(device=device, dtype=torch.float32).reshape(-1, int(np.prod(shape))).contiguous()
            mframes = int(mask.shape[0])
        with torch.cuda.device(device):
            _lib.check(L.pf_cart_create(_lib.ptr(mask), mframes, d, sh, _lib.stream_of(device), ctypes.byref(h)))
        self.handle = h
        self.mframes = mframes
        self.shape = tuple(int(s) for s in shape)
        self.device = device
        self._lib = L

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and self._lib is not None:
            try:
                self._lib.pf_cart_destroy(h)
            except Exception:
                pass
            self.handle = None


@dataclass
class FFTPlan:
    """Same fields as the reference FFTPlan (src/pyfourier/_subroutines/_mask/_plan.py:14-61)."""

    indexes: object  # mask (mframes, *shape) float32, or None
    shape: tuple  # (z, y, x)
    zmap_t_kernel: object  # (L, Nx) complex64 or None
    zmap_s_kernel: object  # (L, *shape) complex64 or None
    zmap_batch_size: object
    device: object = None
    ndim: int = 0
    _native: dict = field(default_factory=dict, repr=False)

    @property
    def mframes(self) -> int:
        return 1 if self.indexes is None else int(self.indexes.shape[0])

    def to(self, device):
        """Place the plan (and its native object) on `device` (a GPU)."""
        dev = resolve_device(device)
        if dev not in self._native:
            self._native[dev] = NativeCart(self.indexes, self.shape, dev)
        if self.indexes is not None:
            self.indexes = self.indexes.to(dev)
        if self.zmap_t_kernel is not None:
            self.zmap_t_kernel = self.zmap_t_kernel.to(dev)
            self.zmap_s_kernel = self.zmap_s_kernel.to(dev)
        self.device = dev
        return self

    def native(self, dev) -> NativeCart:
        self.to(dev)
        return self._native[dev]


def _shape_tuple(shape, ndim):
    if np.isscalar(shape):
        return (int(shape),) * ndim
    return tuple(int(s) for s in shape)


def plan_fft(mask, shape, zmap=None, L=6, nbins=(40, 40), dt=None, T=None, L_batch_size=None) -> FFTPlan:
    """Precompute a Cartesian FFT plan (same signature as the reference plan_fft).

    mask: sampling mask (ncontrasts, [nz], ny, nx) (or without the contrast
    axis), any real dtype, or None; shape: grid size (ndim,) ordered (z, y, x);
    zmap: field map in Hz (real: B0; complex: R2* + i B0) with T (readout
    sampling instants along x, (nx,)) or dt (then T = dt * arange(nx)).
    """
    m = _to_torch(mask)
    ndim = len(shape) if not np.isscalar(shape) else (max(1, m.dim() - 1) if m is not None else 1)
    shp = _shape_tuple(shape, ndim)
    if m is not None:
        if torch.is_complex(m):
            raise ValueError("the k-space mask must be real")
        m = m.to(torch.float32)
        if m.dim() == ndim:
            m = m[None]
        m = m.reshape((-1,) + tuple(m.shape[-ndim:]))
        m = torch.broadcast_to(m, (m.shape[0],) + shp).contiguous()
    bt = bs = None
    if zmap is not None:
        if T is None:
            if dt is None:
                raise ValueError("Please provide raster time dt if T is not known")
            T = dt * torch.arange(shp[-1], dtype=torch.float32)
        T = _to_torch(T).reshape(-1)
        if T.numel() != shp[-1]:
            raise ValueError(f"T must have one entry per readout (x) point, {shp[-1]}")
        bt, bs = mri_exp_approx(_to_torch(zmap), T, L, nbins)
        if tuple(bs.shape[1:]) != shp:
            raise ValueError(f"zmap shape {tuple(bs.shape[1:])} does not match the grid {shp}")
        if L_batch_size is None:
            L_batch_size = int(bt.shape[0])
    dev = m.device if m is not None else None
    return FFTPlan(m, shp, bt, bs, L_batch_size, dev, ndim)


class _CartCall:
    """One operator configuration: plan, basis and B0 segments on a device."""

    def __init__(self, plan: FFTPlan, dev, basis, nframes: int):
        self.plan = plan
        self.native = plan.native(dev)
        self.dev = dev
        self.shape = plan.shape
        self.T = nframes
        self.basis = None
        self.K = 0
        if basis is not None:
            b = _to_torch(basis).to(device=dev, dtype=torch.complex64)
            if b.dim() != 2 or b.shape[0] != nframes:
                raise ValueError(f"basis must have shape (ncontrasts={nframes}, ncoeff), got {tuple(b.shape)}")
            self.basis = b.contiguous()
            self.K = int(b.shape[1])
        self.nimg = self.K if self.K else nframes
        self.segs = [(None, None)]
        if plan.zmap_t_kernel is not None:
            bt = plan.zmap_t_kernel.to(dev)
            bs = plan.zmap_s_kernel.to(dev).reshape(bt.shape[0], -1)
            L = int(bt.shape[0])
            kk = max(1, self.K)
            step = max(1, min(int(plan.zmap_batch_size or L), _RMAX // kk, max(1, 4096 // self.shape[-1])))
            self.segs = [(bt[a : a + step].contiguous(), bs[a : a + step].contiguous()) for a in range(0, L, step)]

    def _desc(self, C, bt, bs):
        d = _lib.BatchDesc()
        d.batch = int(C)
        d.nseg = 1 if bt is None else int(bt.shape[0])
        d.ncoeff = self.K
        d.basis = _lib.ptr(self.basis)
        d.b0_temporal = _lib.ptr(bt)
        d.b0_spatial = _lib.ptr(bs)
        d.nt = self.shape[-1] if bt is not None else 1
        return d

    def run(self, x: torch.Tensor, inverse: bool, code: int) -> torch.Tensor:
        """x (C, nimg, *N) -> (C, T, *N) forward; (C, T, *N) -> (C, nimg, *N) inverse."""
        C = int(x.shape[0])
        nout = self.nimg if inverse else self.T
        out = torch.empty((C, nout) + self.shape, dtype=torch.complex64, device=self.dev)
        if C == 0:
            return out
        L = self.native._lib
        fn = L.pf_cart_ifft if inverse else L.pf_cart_fft
        stream = _lib.stream_of(self.dev)
        for i, (bt, bs) in enumerate(self.segs):
            dst = out if i == 0 else torch.empty_like(out)
            desc = self._desc(C, bt, bs)
            n = ctypes.c_size_t()
            _lib.check(L.pf_cart_workspace_size(self.native.handle, ctypes.byref(desc), self.T, ctypes.byref(n)))
            ws = torch.empty(max(1, n.value), dtype=torch.uint8, device=self.dev)
            with torch.cuda.device(self.dev):
                _lib.check(fn(self.native.handle, ctypes.byref(desc), self.T, _lib.ptr(x), _lib.ptr(dst), int(code),
                              _lib.ptr(ws), n.value, stream))
            if i:
                out += dst
        return out


class _CartFn(torch.autograd.Function):
    """Autograd: the backward of either direction is its exact adjoint."""

    @staticmethod
    def forward(ctx, x, call, inverse, code):
        ctx.call, ctx.inverse, ctx.code = call, inverse, code
        return call.run(x, inverse, code)

    @staticmethod
    def backward(ctx, g):
        gx = ctx.call.run(g.contiguous().to(torch.complex64), not ctx.inverse, _adjoint_code(ctx.code))
        return gx, None, None, None


def _layout(x: torch.Tensor, ndim: int, plan: FFTPlan, basis, inverse: bool):
    """(batch shape, C, nframes, planes of the input) for the reference shape rules."""
    if tuple(x.shape[-ndim:]) != tuple(plan.shape):
        raise ValueError(f"spatial shape {tuple(x.shape[-ndim:])} does not match the plan's {tuple(plan.shape)}")
    mf = plan.mframes
    if basis is not None:
        b = _to_torch(basis)
        T, K = int(b.shape[0]), int(b.shape[1])
        if mf not in (1, T):
            raise ValueError(f"mask has {mf} frames, basis {T} contrasts")
        planes = T if inverse else K
        if x.dim() < ndim + 1 or x.shape[-ndim - 1] != planes:
            raise ValueError(f"expected {planes} planes before the spatial axes, got shape {tuple(x.shape)}")
        batch = tuple(x.shape[: -ndim - 1])
        return batch, T, planes, True
    if mf > 1:
        if x.dim() < ndim + 1 or x.shape[-ndim - 1] != mf:
            raise ValueError(f"expected {mf} contrasts before the spatial axes, got shape {tuple(x.shape)}")
        return tuple(x.shape[: -ndim - 1]), mf, mf, True
    return tuple(x.shape[:-ndim]), 1, 1, False


def _apply(ndim, data, mask, plan, basis, zmap, L, nbins, dt, T, L_batch_size, norm, device, inverse):
    code = _norm_code(norm)
    backend_numpy = not isinstance(data, torch.Tensor)
    x = _to_torch(data)
    idevice = x.device
    dev = resolve_device(idevice if device is None else device)
    if plan is None:
        plan = plan_fft(mask, tuple(x.shape[-ndim:]), zmap, L, nbins, dt, T, L_batch_size)
    plan.to(dev)
    dtype = x.dtype
    batch, nfr, planes, has_axis = _layout(x, ndim, plan, basis, inverse)
    C = int(np.prod(batch)) if batch else 1
    call = _CartCall(plan, dev, basis, nfr)
    xin = x.to(device=dev, dtype=torch.complex64).reshape((C, planes) + tuple(plan.shape)).contiguous()
    y = _CartFn.apply(xin, call, inverse, code)
    nout = call.nimg if inverse else nfr
    y = y.reshape(batch + ((nout,) if has_axis else ()) + tuple(plan.shape))
    return _restore(y, dtype if dtype.is_floating_point or dtype.is_complex else torch.complex64, backend_numpy,
                    idevice)


def fftn(ndim, image, mask=None, plan=None, basis=None, zmap=None, L=6, nbins=(40, 40), dt=None, T=None,
         L_batch_size=None, norm=None, device="cpu"):
    """N-dimensional masked centred FFT (drop-in for pyfourier.fftn).

    image (..., [ncontrasts | ncoeff], [nz], ny, nx) -> kspace (..., [ncontrasts], [nz], ny, nx);
    `device` must name a GPU.
    """
    return _apply(ndim, image, mask, plan, basis, zmap, L, nbins, dt, T, L_batch_size, norm, device, False)


def ifftn(ndim, kspace, mask=None, plan=None, basis=None, zmap=None, L=6, nbins=(40, 40), dt=None, T=None,
          L_batch_size=None, norm=None, device="cpu", verbose=0):
    """N-dimensional masked centred inverse FFT (drop-in for pyfourier.ifftn).

    kspace (..., [ncontrasts], [nz], ny, nx) -> image (..., [ncontrasts | ncoeff], [nz], ny, nx).
    """
    return _apply(ndim, kspace, mask, plan, basis, zmap, L, nbins, dt, T, L_batch_size, norm, device, True)


# ---------------------------------------------------------------------------
# centred FFT utilities: src/pyfourier/_subroutines/_fft.py:11-194
# ---------------------------------------------------------------------------
_UTIL_PLANS: dict = {}
_UTIL_CODES = {  # (forward code, inverse code) of torch.fft's norm names
    None: (_lib.PF_NORM_NONE, _lib.PF_NORM_NONE),
    "backward": (_lib.PF_NORM_NONE, _lib.PF_NORM_NONE),
    "ortho": (_lib.PF_NORM_ORTHO, _lib.PF_NORM_ORTHO),
    "forward": (_lib.PF_NORM_ADJOINT, _lib.PF_NORM_ADJOINT),
}


def _util_plan(shape, dev) -> FFTPlan:
    key = (tuple(shape), dev)
    p = _UTIL_PLANS.get(key)
    if p is None:
        p = FFTPlan(None, tuple(shape), None, None, None, None, len(shape))
        p.to(dev)
        _UTIL_PLANS[key] = p
    return p


def _fft_util(data, axes, norm, centered, inverse):
    if norm not in _UTIL_CODES:
        raise ValueError(f"norm must be None, 'backward', 'ortho' or 'forward', got {norm!r}")
    backend_numpy = not isinstance(data, torch.Tensor)
    x = _to_torch(data)
    idevice = x.device
    dev = resolve_device(idevice if x.is_cuda else torch.device("cuda", torch.cuda.current_device()))
    nd = x.dim()
    ax = list(range(nd)) if axes is None else sorted({int(a) % nd for a in np.atleast_1d(axes)})
    if not 1 <= len(ax) <= 3:
        raise ValueError("the native FFT transforms 1 to 3 axes")
    d = len(ax)
    dst = list(range(nd - d, nd))
    xm = torch.movedim(x, ax, dst).to(device=dev, dtype=torch.complex64)
    shape = tuple(xm.shape[-d:])
    sax = tuple(range(-d, 0))
    if not centered:
        xm = torch.fft.fftshift(xm, dim=sax)
    bshape = tuple(xm.shape[:-d])
    C = int(np.prod(bshape)) if bshape else 1
    call = _CartCall(_util_plan(shape, dev), dev, None, 1)
    code = _UTIL_CODES[norm][1 if inverse else 0]
    y = _CartFn.apply(xm.reshape((C, 1) + shape).contiguous(), call, inverse, code).reshape(bshape + shape)
    if not centered:
        y = torch.fft.ifftshift(y, dim=sax)
    y = torch.movedim(y, dst, ax)
    if x.dtype in (torch.float64, torch.complex128):
        y = y.to(torch.complex128)
    if backend_numpy:
        return y.detach().cpu().numpy()
    return y.to(idevice)


def fft(input, axes=None, norm="ortho", centered=True):
    """Centred FFT over `axes` (drop-in for pyfourier.fft); runs on the GPU."""
    return _fft_util(input, axes, norm, centered, False)


def ifft(input, axes=None, norm="ortho", centered=True):
    """Centred inverse FFT over `axes` (drop-in for pyfourier.ifft); runs on the GPU."""
    return _fft_util(input, axes, norm, centered, True)
