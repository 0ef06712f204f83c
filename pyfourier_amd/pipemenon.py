"""Pipe-Menon density compensation: drop-in pipemenon.pipe_menon_dcf
(the reference exports this module as pyfourier.pipemenon,
src/pyfourier/__init__.py:15).

Mirrors src/pyfourier/_dcomp/pipemenon.py:9-115: starting from w = 1,
iterate w <- w / |G G^H w| with G the forward NUFFT under norm="ortho"
(width 6, oversampling 1.5 by default), where, as in the reference, each
operator's output is cast back to the real dtype of w (the real part is
kept).  Here one iteration (adjoint, real part, forward, real part, divide)
is recorded once as a HIP graph and replayed `niter` times on static
buffers, so the whole estimate is `niter` graph launches.
"""

from __future__ import annotations

import torch

from . import _lib
from ._nufft import _Call, _to_torch
from ._plan import plan_nufft, resolve_device


def pipe_menon_dcf(coord, shape, niter=30, threadsperblock=128, device=None, width=6, oversamp=1.5, graph=True):
    """Density compensation weights of shape coord.shape[:-1] (float32).

    `device` must name a GPU (defaults to coord's device when that is one,
    else cuda:0); `threadsperblock` is accepted and ignored.
    """
    c = _to_torch(coord)
    backend_numpy = not isinstance(coord, torch.Tensor)
    idevice = c.device
    if device is None:
        device = idevice if idevice.type == "cuda" else "cuda"
    dev = resolve_device(device)
    plan = plan_nufft(c.detach().to(torch.float32), shape, width, oversamp).to(dev)
    call = _Call(plan, dev, None, None, "ortho")
    it = call.interp
    code = _lib.PF_NORM_ORTHO
    w = torch.ones((1, it.nframes, it.npts), dtype=torch.complex64, device=dev)
    img = torch.empty((1, call.nimg) + tuple(it.shape), dtype=torch.complex64, device=dev)
    ghg = torch.empty_like(w)

    def step():
        call.adjoint(w, code, out=img)
        img.imag.zero_()
        call.forward(img, code, out=ghg)
        w.real.div_(ghg.real.abs())

    with torch.cuda.device(dev):
        if graph and niter > 1:
            side = torch.cuda.Stream(dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                step()  # warm-up on a side stream; counts as iteration 1
            torch.cuda.current_stream(dev).wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                step()
            # capture records without executing; run the remaining iterations
            for _ in range(niter - 1):
                g.replay()
        else:
            for _ in range(niter):
                step()
    out = w.real.reshape(c.shape[:-1]).contiguous()
    if backend_numpy:
        return out.cpu().numpy()
    return out.to(idevice)
