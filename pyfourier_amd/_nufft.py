"""Forward and adjoint NUFFT: drop-in nufft / nufft_adj.

Mirrors src/pyfourier/_nufft/_fwd.py:18-230, src/pyfourier/_nufft/_adj.py:18-216
and the operator pipeline of src/pyfourier/_nufft/_nufft.py:11-287.  Every
call goes through the native engine (libpfhip, include/pfhip.h); there is no
CPU or PyTorch fallback.

Deviations from the reference (its defects, see the module docstrings of
_plan/_traj/_b0): B0 segments are summed once (A.13); autograd backward
passes use the exact adjoint of the forward in use, also under norm=None
(A.22); inputs are never modified in place (A.19).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._plan import NUFFTPlan, plan_nufft, resolve_device

_NORMS = {None: _lib.PF_NORM_NONE, "ortho": _lib.PF_NORM_ORTHO}


class stage_timer:
    """Context manager that runs operator calls stage by stage (the same
    native launchers as the fused entry points) with HIP events recorded on
    the call's stream around every stage.

    >>> with stage_timer() as t:
    ...     nufft(...)
    >>> t.summary()  # {stage: [ms per call, ...]} after a synchronize
    """

    active = None

    def __init__(self):
        self.records = []

    def __enter__(self):
        stage_timer.active = self
        return self

    def __exit__(self, *exc):
        stage_timer.active = None
        return False

    def record(self, name, fn, device):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        st = torch.cuda.current_stream(device)
        s.record(st)
        fn()
        e.record(st)
        self.records.append((name, s, e))

    def summary(self):
        out = {}
        for name, s, e in self.records:
            out.setdefault(name, []).append(s.elapsed_time(e))
        return out


def _norm_code(norm) -> int:
    if norm not in _NORMS:
        raise ValueError(f"norm must be None or 'ortho', got {norm!r}")
    return _NORMS[norm]


def _to_torch(x):
    if x is None or isinstance(x, torch.Tensor):
        return x
    return torch.as_tensor(np.asarray(x))


class _Call:
    """Everything an operator call needs besides the data tensors."""

    def __init__(self, plan: NUFFTPlan, device, basis, weight, norm):
        self.plan = plan
        self.device = device
        self.interp = plan.interpolator
        self.native = self.interp.native(device)
        self.norm = norm
        self.basis = None
        if basis is not None:
            b = _to_torch(basis)
            if b.dim() != 2 or b.shape[0] != self.interp.nframes:
                raise ValueError(
                    f"basis must have shape (ncontrasts={self.interp.nframes}, ncoeff), got {tuple(b.shape)}"
                )
            self.basis = b.to(device=device, dtype=torch.complex64).contiguous()
        self.weight_flat = None  # (F*npts,) float32 for the native path
        self.weight_full = None  # anything else, applied with torch
        if weight is not None:
            w = _to_torch(weight).to(device)
            if not torch.is_complex(w):
                ks = (self.interp.nframes,) + tuple(self.interp.dshape)
                try:
                    we = torch.broadcast_to(w.to(torch.float32), ks)
                    if we.dim() >= w.dim():
                        self.weight_flat = we.reshape(-1).contiguous()
                except RuntimeError:
                    self.weight_flat = None
            if self.weight_flat is None:
                self.weight_full = w.to(torch.complex64 if torch.is_complex(w) else torch.float32)
        self.segments = self._segments()

    def _segments(self):
        p = self.plan
        if p.zmap_s_kernel is None:
            return [(None, None)]
        bt = p.zmap_t_kernel.to(device=self.device, dtype=torch.complex64)
        bs = p.zmap_s_kernel.to(device=self.device, dtype=torch.complex64)
        L = bt.shape[0]
        step = int(p.zmap_batch_size or L)
        out = []
        for l0 in range(0, L, step):
            l1 = min(L, l0 + step)
            out.append((bt[l0:l1].contiguous(), bs[l0:l1].reshape(l1 - l0, -1).contiguous()))
        return out

    @property
    def nimg(self):
        return self.basis.shape[1] if self.basis is not None else self.interp.nframes

    def desc(self, C, seg):
        d = _lib.BatchDesc()
        d.batch = int(C)
        d.ncoeff = int(self.basis.shape[1]) if self.basis is not None else 0
        d.basis = _lib.ptr(self.basis)
        bt, bs = seg
        d.nseg = int(bt.shape[0]) if bt is not None else 1
        d.b0_temporal = _lib.ptr(bt)
        d.b0_spatial = _lib.ptr(bs)
        d.nt = int(bt.shape[1]) if bt is not None else 1
        return d

    def describe(self, C: int, seg=(None, None)) -> str:
        """Names of the gather / spread kernel instances a batch of C runs
        (pf_describe_launch), e.g. for tests and profiles."""
        buf = ctypes.create_string_buffer(512)
        desc = self.desc(C, seg)
        _lib.check(_lib.lib().pf_describe_launch(self.native.handle, ctypes.byref(desc), buf, 512))
        return buf.value.decode()

    # ---- native operator calls on (C, nimg, *N) / (C, F, npts) ------------
    def forward(self, x: torch.Tensor, norm_code: int, out: torch.Tensor | None = None) -> torch.Tensor:
        C = x.shape[0]
        it = self.interp
        if out is None:
            out = torch.empty((C, it.nframes, it.npts), dtype=torch.complex64, device=self.device)
        if C == 0:
            return out
        L = _lib.lib()
        acc = None
        for seg in self.segments:
            desc = self.desc(C, seg)
            nbytes = ctypes.c_size_t()
            _lib.check(L.pf_workspace_size(self.native.handle, ctypes.byref(desc), 0, ctypes.byref(nbytes)))
            ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.device)
            y = out if acc is None else torch.empty_like(out)
            with torch.cuda.device(self.device):
                if stage_timer.active is not None:
                    self._staged_forward(desc, x, y, ws, norm_code)
                else:
                    _lib.check(
                        L.pf_nufft(
                            self.native.handle, ctypes.byref(desc), _lib.ptr(x), _lib.ptr(y),
                            _lib.ptr(self.weight_flat), norm_code, _lib.ptr(ws), nbytes.value,
                            _lib.stream_of(self.device),
                        )
                    )
            acc = y if acc is None else acc.add_(y)
        return acc

    # ---- forward that keeps its grids (for the trajectory gradient) ----------
    def forward_keep(self, x: torch.Tensor, norm_code: int):
        """forward(x) that also returns, per B0 segment chunk, the workspace
        holding the oversampled k-space grid of x, so the trajectory gradient
        does not transform x again."""
        C = x.shape[0]
        it = self.interp
        out = torch.empty((C, it.nframes, it.npts), dtype=torch.complex64, device=self.device)
        L = _lib.lib()
        kept, acc = [], None
        fac = self._factor(norm_code, False)
        for seg in self.segments:
            desc = self.desc(C, seg)
            nbytes = ctypes.c_size_t()
            _lib.check(L.pf_workspace_size(self.native.handle, ctypes.byref(desc), 1, ctypes.byref(nbytes)))
            ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.device)
            grid, ys, fws, _ = self._split_ws(desc, ws)
            y = out if acc is None else torch.empty_like(out)
            h, d, st = self.native.handle, ctypes.byref(desc), _lib.stream_of(self.device)
            with torch.cuda.device(self.device):
                _lib.check(L.pf_pad_fft(h, d, _lib.ptr(x), _lib.ptr(grid), _lib.ptr(fws), st))
                _lib.check(L.pf_interp(h, d, _lib.ptr(grid), _lib.ptr(ys), st))
                _lib.check(L.pf_unsort_samples(h, d, _lib.ptr(ys), _lib.ptr(y), _lib.ptr(self.weight_flat), fac, st))
            kept.append(ws)
            acc = y if acc is None else acc.add_(y)
        return acc, kept

    # ---- coordinate gradient ------------------------------------------------
    def coord_grad(self, img, h: torch.Tensor, factor: float, kept=None) -> torch.Tensor:
        """Re(sum_c conj(h) * d(A img)/d coord) per sample, in user units.

        img (C, nimg, *N) is the image whose forward grid is differentiated
        (or None with `kept`, the workspaces forward_keep returned for it);
        h (C, F, npts) the paired k-space (weight applied here); factor the
        operator's 1/scale and norm factor.  Returns (F*npts, d) ordered
        (x, y, z), written by the native kernel at the original sample order.
        """
        it = self.interp
        C = h.shape[0]
        M = it.nframes * it.npts
        dd = it.ndim
        L = _lib.lib()
        # d c_os / d coord per stored axis (z, y, x): G/N times the unit factor
        sc = [1.0, 1.0, 1.0]
        for a in range(dd):
            sc[3 - dd + a] = factor * it.os_shape[a] / it.shape[a] * it.unit_scale[dd - 1 - a]
        scale = (ctypes.c_float * 3)(*sc)
        out = torch.empty((M, dd), dtype=torch.float32, device=self.device)
        for i, seg in enumerate(self.segments):
            desc = self.desc(C, seg)
            if kept is not None:
                ws = kept[i]
            else:
                nbytes = ctypes.c_size_t()
                _lib.check(L.pf_workspace_size(self.native.handle, ctypes.byref(desc), 1, ctypes.byref(nbytes)))
                ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.device)
            grid, ys, fws, _ = self._split_ws(desc, ws)
            h_, d, st = self.native.handle, ctypes.byref(desc), _lib.stream_of(self.device)
            with torch.cuda.device(self.device):
                if kept is None:
                    _lib.check(L.pf_pad_fft(h_, d, _lib.ptr(img), _lib.ptr(grid), _lib.ptr(fws), st))
                _lib.check(L.pf_sort_samples(h_, d, _lib.ptr(h), _lib.ptr(ys), _lib.ptr(self.weight_flat), st))
                _lib.check(L.pf_interp_grad(h_, d, _lib.ptr(grid), _lib.ptr(ys), scale, int(i > 0), _lib.ptr(out), st))
        return out

    # ---- stage-by-stage execution (timing) ---------------------------------
    def _factor(self, norm_code: int, adjoint: bool) -> float:
        g = float(np.prod(self.interp.os_shape))
        f = 1.0 / self.native.info.scale
        if norm_code == _lib.PF_NORM_ORTHO:
            return f / math.sqrt(g)
        if (norm_code == _lib.PF_NORM_NONE) == adjoint:
            return f / g
        return f

    def _split_ws(self, desc, ws):
        L = _lib.lib()
        off = [ctypes.c_size_t() for _ in range(4)]
        _lib.check(L.pf_workspace_offsets(self.native.handle, ctypes.byref(desc), *[ctypes.byref(o) for o in off]))
        g, s, f, p = (o.value for o in off)
        return ws[g:s], ws[s:f], ws[f:p], ws[p:]

    def _staged_forward(self, desc, x, y, ws, norm_code):
        L = _lib.lib()
        h, d, st = self.native.handle, ctypes.byref(desc), _lib.stream_of(self.device)
        grid, ys, fws, _ = self._split_ws(desc, ws)
        t = stage_timer.active
        dev = self.device
        t.record("pad_fft", lambda: _lib.check(
            L.pf_pad_fft(h, d, _lib.ptr(x), _lib.ptr(grid), _lib.ptr(fws), st)), dev)
        t.record("interp", lambda: _lib.check(L.pf_interp(h, d, _lib.ptr(grid), _lib.ptr(ys), st)), dev)
        fac = self._factor(norm_code, False)
        t.record("unsort", lambda: _lib.check(
            L.pf_unsort_samples(h, d, _lib.ptr(ys), _lib.ptr(y), _lib.ptr(self.weight_flat), fac, st)), dev)

    def _staged_adjoint(self, desc, y, x, ws, norm_code):
        L = _lib.lib()
        h, d, st = self.native.handle, ctypes.byref(desc), _lib.stream_of(self.device)
        grid, ys, fws, part = self._split_ws(desc, ws)
        t = stage_timer.active
        dev = self.device
        t.record("sort", lambda: _lib.check(
            L.pf_sort_samples(h, d, _lib.ptr(y), _lib.ptr(ys), _lib.ptr(self.weight_flat), st)), dev)
        t.record("spread", lambda: _lib.check(
            L.pf_spread(h, d, _lib.ptr(ys), _lib.ptr(grid), _lib.ptr(part), st)), dev)
        fac = self._factor(norm_code, True)
        t.record("ifft_crop", lambda: _lib.check(
            L.pf_ifft_crop(h, d, _lib.ptr(grid), _lib.ptr(x), fac, _lib.ptr(fws), st)), dev)

    def adjoint(self, y: torch.Tensor, norm_code: int, out: torch.Tensor | None = None) -> torch.Tensor:
        C = y.shape[0]
        it = self.interp
        if out is None:
            out = torch.empty((C, self.nimg) + tuple(it.shape), dtype=torch.complex64, device=self.device)
        if C == 0:
            return out
        L = _lib.lib()
        acc = None
        for seg in self.segments:
            desc = self.desc(C, seg)
            nbytes = ctypes.c_size_t()
            _lib.check(L.pf_workspace_size(self.native.handle, ctypes.byref(desc), 1, ctypes.byref(nbytes)))
            ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.device)
            x = out if acc is None else torch.empty_like(out)
            with torch.cuda.device(self.device):
                if stage_timer.active is not None:
                    self._staged_adjoint(desc, y, x, ws, norm_code)
                else:
                    _lib.check(
                        L.pf_nufft_adj(
                            self.native.handle, ctypes.byref(desc), _lib.ptr(y), _lib.ptr(x),
                            _lib.ptr(self.weight_flat), norm_code, _lib.ptr(ws), nbytes.value,
                            _lib.stream_of(self.device),
                        )
                    )
            acc = x if acc is None else acc.add_(x)
        return acc


# ---- shape handling (src/pyfourier/_subroutines/_utils/_common.py:53-71, intent) ----
def _image_layout(image: torch.Tensor, call: _Call):
    it = call.interp
    d = it.ndim
    if tuple(image.shape[-d:]) != tuple(it.shape):
        raise ValueError(f"image spatial shape {tuple(image.shape[-d:])} does not match plan shape {it.shape}")
    if call.basis is not None or it.nframes > 1:
        nimg = call.nimg
        if image.dim() < d + 1 or image.shape[-(d + 1)] != nimg:
            what = "ncoeff" if call.basis is not None else "ncontrasts"
            raise ValueError(f"image must have a {what}={nimg} axis before the spatial axes")
        batch_shape = tuple(image.shape[: -(d + 1)])
    else:
        nimg = 1
        batch_shape = tuple(image.shape[:-d])
    C = int(np.prod(batch_shape)) if batch_shape else 1
    return batch_shape, C, nimg


def _kspace_layout(kspace: torch.Tensor, call: _Call):
    it = call.interp
    nd = len(it.dshape)
    if tuple(kspace.shape[-nd:]) != tuple(it.dshape):
        raise ValueError(f"k-space shape {tuple(kspace.shape)} does not match the trajectory {it.dshape}")
    if it.nframes > 1:
        if kspace.shape[-(nd + 1)] != it.nframes:
            raise ValueError(f"k-space must have an ncontrasts={it.nframes} axis")
        batch_shape = tuple(kspace.shape[: -(nd + 1)])
    else:
        batch_shape = tuple(kspace.shape[:-nd])
    C = int(np.prod(batch_shape)) if batch_shape else 1
    return batch_shape, C


# ---- autograd --------------------------------------------------------------
class _NUFFTFn(torch.autograd.Function):
    """y = A x with A the forward NUFFT; backward is the exact adjoint."""

    @staticmethod
    def forward(ctx, image, call, norm_code):
        ctx.call = call
        ctx.norm_code = norm_code
        return call.forward(image, norm_code)

    @staticmethod
    def backward(ctx, g):
        return ctx.call.adjoint(g.contiguous().to(torch.complex64), _adjoint_code(ctx.norm_code)), None, None


class _NUFFTAdjFn(torch.autograd.Function):
    """x = A^H y with the adjoint NUFFT; backward is the exact adjoint (A)."""

    @staticmethod
    def forward(ctx, kspace, call, norm_code):
        ctx.call = call
        ctx.norm_code = norm_code
        return call.adjoint(kspace, norm_code)

    @staticmethod
    def backward(ctx, g):
        return ctx.call.forward(g.contiguous().to(torch.complex64), _adjoint_code(ctx.norm_code)), None, None


def _adjoint_code(norm_code: int) -> int:
    if norm_code == _lib.PF_NORM_ORTHO:
        return _lib.PF_NORM_ORTHO
    return _lib.PF_NORM_ADJOINT if norm_code == _lib.PF_NORM_NONE else _lib.PF_NORM_NONE


def _coord_grad(call: _Call, coord: torch.Tensor, img, h: torch.Tensor, adjoint: bool, norm_code: int, kept=None):
    g = call.coord_grad(img, h, call._factor(norm_code, adjoint), kept)
    return g.reshape(coord.shape).to(device=coord.device, dtype=coord.dtype)


class _NUFFTTrajFn(torch.autograd.Function):
    """y = A(coord) x with gradients for the image and the trajectory.

    Replaces NUFFTTraj of src/pyfourier/_nufft/_fwd.py:275-319: the image
    gradient is the exact adjoint, the coordinate gradient
    Re(conj(g) dy/dcoord) comes from the Kaiser-Bessel-derivative gather.
    """

    @staticmethod
    def forward(ctx, image, coord, call, norm_code):
        ctx.call = call
        ctx.norm_code = norm_code
        ctx.save_for_backward(coord)
        # keep the forward's k-space grid: the coordinate gradient reads it
        y, ctx.kept = call.forward_keep(image, norm_code)
        return y

    @staticmethod
    def backward(ctx, g):
        call = ctx.call
        (coord,) = ctx.saved_tensors
        g = g.contiguous().to(torch.complex64)
        gi = call.adjoint(g, _adjoint_code(ctx.norm_code)) if ctx.needs_input_grad[0] else None
        gc = _coord_grad(call, coord, None, g, False, ctx.norm_code, ctx.kept) if ctx.needs_input_grad[1] else None
        ctx.kept = None
        return gi, gc, None, None


class _NUFFTAdjTrajFn(torch.autograd.Function):
    """x = A(coord)^H y with gradients for k-space and the trajectory
    (replaces NUFFTAdjointTraj of src/pyfourier/_nufft/_adj.py:261-306)."""

    @staticmethod
    def forward(ctx, kspace, coord, call, norm_code):
        ctx.call = call
        ctx.norm_code = norm_code
        ctx.save_for_backward(kspace, coord)
        return call.adjoint(kspace, norm_code)

    @staticmethod
    def backward(ctx, gx):
        call = ctx.call
        kspace, coord = ctx.saved_tensors
        gx = gx.contiguous().to(torch.complex64)
        gk = call.forward(gx, _adjoint_code(ctx.norm_code)) if ctx.needs_input_grad[0] else None
        gc = _coord_grad(call, coord, gx, kspace, True, ctx.norm_code) if ctx.needs_input_grad[1] else None
        return gk, gc, None, None


def _restore(out: torch.Tensor, dtype, backend_numpy: bool, idevice):
    if not dtype.is_complex:
        out = out.real
    out = out.to(dtype)
    if backend_numpy:
        return out.detach().cpu().numpy()
    return out.to(idevice)


def _traj_tensor(coord, call):
    """The coordinate tensor to differentiate, when the caller asks for it.

    As in the reference, the gradient is taken at the plan's trajectory (the
    one ``coord`` built, or the one a given ``nufft_plan`` was built from).
    """
    if isinstance(coord, torch.Tensor) and coord.requires_grad and torch.is_grad_enabled():
        it = call.interp
        if coord.numel() != it.nframes * it.npts * it.ndim:
            raise ValueError(f"coord {tuple(coord.shape)} does not match the plan's trajectory")
        return coord
    return None


def _prepare(x, coord, plan, shape_arg, zmap, L, nbins, dt, T, L_batch_size, width, oversamp, device, for_adj,
             units="auto"):
    backend_numpy = not isinstance(x, torch.Tensor)
    x = _to_torch(x)
    idevice = x.device
    dev = resolve_device(idevice if device is None else device)
    if plan is None:
        if coord is None:
            raise ValueError("provide coord or nufft_plan")
        c = _to_torch(coord)
        if shape_arg is None:
            if for_adj:
                raise ValueError("nufft_adj needs the image shape when no plan is given")
            shape_arg = tuple(x.shape[-c.shape[-1]:])
        plan = plan_nufft(c.detach().to(torch.float32), shape_arg, width, oversamp, zmap, L, nbins, dt, T, L_batch_size,
                          units=units)
    plan.to(dev)
    return x, backend_numpy, idevice, dev, plan


def nufft(
    image,
    coord=None,
    shape=None,
    nufft_plan=None,
    basis=None,
    zmap=None,
    L=6,
    nbins=(40, 40),
    dt=None,
    T=None,
    L_batch_size=None,
    weight=None,
    norm=None,
    device="cpu",
    threadsperblock=128,
    width=4,
    oversamp=1.25,
    units="auto",
):
    """N-dimensional forward NUFFT (drop-in for pyfourier.nufft).

    image (..., [ncontrasts | ncoeff], [nz], ny, nx) -> kspace (..., [ncontrasts], nviews, nsamples).
    `device` must name a GPU ("cuda:n" or n); `threadsperblock` is accepted and ignored.
    `units` (extension): unit of coord when no plan is given, as for plan_nufft.
    """
    code = _norm_code(norm)
    x, backend_numpy, idevice, dev, plan = _prepare(
        image, coord, nufft_plan, shape, zmap, L, nbins, dt, T, L_batch_size, width, oversamp, device, False, units
    )
    dtype = x.dtype
    call = _Call(plan, dev, basis, weight, norm)
    batch_shape, C, nimg = _image_layout(x, call)
    xin = x.to(device=dev, dtype=torch.complex64).reshape((C, nimg) + tuple(call.interp.shape)).contiguous()
    tc = _traj_tensor(coord, call)
    y = _NUFFTFn.apply(xin, call, code) if tc is None else _NUFFTTrajFn.apply(xin, tc, call, code)
    it = call.interp
    oshape = batch_shape + ((it.nframes,) if it.nframes > 1 else ()) + tuple(it.dshape)
    y = y.reshape(oshape)
    if call.weight_full is not None:
        y = y * call.weight_full
    return _restore(y, dtype if dtype.is_floating_point or dtype.is_complex else torch.complex64, backend_numpy, idevice)


def nufft_adj(
    kspace,
    coord=None,
    shape=None,
    nufft_plan=None,
    basis=None,
    zmap=None,
    L=6,
    nbins=(40, 40),
    dt=None,
    T=None,
    L_batch_size=None,
    weight=None,
    norm=None,
    device="cpu",
    threadsperblock=128,
    width=4,
    oversamp=1.25,
    units="auto",
):
    """N-dimensional adjoint NUFFT (drop-in for pyfourier.nufft_adj).

    kspace (..., [ncontrasts], nviews, nsamples) -> image (..., [ncontrasts | ncoeff], [nz], ny, nx).
    `units` (extension): unit of coord when no plan is given, as for plan_nufft.
    """
    code = _norm_code(norm)
    y, backend_numpy, idevice, dev, plan = _prepare(
        kspace, coord, nufft_plan, shape, zmap, L, nbins, dt, T, L_batch_size, width, oversamp, device, True, units
    )
    dtype = y.dtype
    call = _Call(plan, dev, basis, weight, norm)
    batch_shape, C = _kspace_layout(y, call)
    it = call.interp
    yin = y.to(device=dev, dtype=torch.complex64)
    if call.weight_full is not None:
        yin = yin * call.weight_full
    yin = yin.reshape(C, it.nframes, it.npts).contiguous()
    tc = _traj_tensor(coord, call)
    x = _NUFFTAdjFn.apply(yin, call, code) if tc is None else _NUFFTAdjTrajFn.apply(yin, tc, call, code)
    keep = call.basis is not None or it.nframes > 1
    oshape = batch_shape + ((call.nimg,) if keep else ()) + tuple(it.shape)
    x = x.reshape(oshape)
    return _restore(x, dtype if dtype.is_floating_point or dtype.is_complex else torch.complex64, backend_numpy, idevice)
