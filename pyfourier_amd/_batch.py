"""Uneven batches of independent NUFFT problems replayed as one HIP graph.

A reconstruction or training loop over many slices, each with its own
matrix size, trajectory and sample count, pays the per-call host cost of the
drop-in API (argument checks, plan lookup, workspace allocation, one ctypes
round trip per native call) once per slice.  NUFFTBatch plans every item
once, owns static image and k-space buffers for all items, records every
item's native launches into one torch.cuda.CUDAGraph (a hipGraph on ROCm)
per direction and norm, and replays it: one host call for the whole uneven
batch.  Each item keeps its own plan; the items run back to back on one
stream inside the graph.

No counterpart in the reference (which loops in Python); the operator of
each item is exactly the one nufft / nufft_adj apply with the same plan.
"""

from __future__ import annotations

import torch

from . import _lib
from ._nufft import _Call, _adjoint_code, _norm_code, _to_torch
from ._plan import plan_nufft, resolve_device


class NUFFTBatch:
    """Forward / adjoint NUFFT of an uneven list of problems in one launch.

    coords[i]: trajectory of item i, (ncontrasts, nviews, nsamples, ndim) or
    fewer leading axes, in any unit; shapes[i]: its image matrix.  Every
    item carries `nbatch` batch entries (coils).  weights[i] (optional) is a
    real density weight broadcastable to item i's k-space; basis (optional,
    shared) is a (ncontrasts, ncoeff) subspace basis.

    ``images[i]`` (nbatch, nimg_i, *shape_i) and ``kspace[i]`` (nbatch,
    ncontrasts_i, npts_i) are the static buffers the graphs read and write:
    fill them in place (or pass lists to forward/adjoint, which copy with one
    multi-tensor copy) and read the results from the other side.  A replay
    overwrites the previous results.
    """

    def __init__(self, coords, shapes, nbatch=1, width=4, oversamp=1.25, weights=None, basis=None, norm=None,
                 device="cuda", graph=True):
        if len(coords) != len(shapes):
            raise ValueError("coords and shapes must have the same length")
        if len(coords) == 0:
            raise ValueError("empty batch")
        self.device = resolve_device(device)
        self.norm_code = _norm_code(norm)
        self.use_graph = bool(graph)
        self.calls = []
        self.images = []
        self.kspace = []
        for i, (c, sh) in enumerate(zip(coords, shapes)):
            plan = plan_nufft(_to_torch(c).to(torch.float32), sh, width, oversamp).to(self.device)
            w = None if weights is None else weights[i]
            call = _Call(plan, self.device, basis, w, norm)
            if call.weight_full is not None:
                raise ValueError("NUFFTBatch takes real density weights broadcastable to each item's k-space")
            it = call.interp
            self.calls.append(call)
            self.images.append(torch.zeros((nbatch, call.nimg) + tuple(it.shape), dtype=torch.complex64,
                                           device=self.device))
            self.kspace.append(torch.zeros((nbatch, it.nframes, it.npts), dtype=torch.complex64,
                                           device=self.device))
        self._graphs = {}

    def __len__(self):
        return len(self.calls)

    # ---- one pass over all items, eager ----------------------------------
    def _launch(self, adjoint: bool, code: int):
        for call, x, y in zip(self.calls, self.images, self.kspace):
            if adjoint:
                call.adjoint(y, code, out=x)
            else:
                call.forward(x, code, out=y)

    def _replay(self, adjoint: bool, code: int):
        if not self.use_graph:
            with torch.cuda.device(self.device):
                self._launch(adjoint, code)
            return
        key = (adjoint, code)
        g = self._graphs.get(key)
        if g is None:
            with torch.cuda.device(self.device):
                # warm-up outside the capture: lazily built FFT state and
                # the caching allocator's blocks exist before recording
                side = torch.cuda.Stream(self.device)
                side.wait_stream(torch.cuda.current_stream(self.device))
                with torch.cuda.stream(side):
                    self._launch(adjoint, code)
                torch.cuda.current_stream(self.device).wait_stream(side)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._launch(adjoint, code)
            self._graphs[key] = g
        g.replay()

    def forward(self, images=None):
        """kspace[i] = nufft(images[i]) for every item; returns self.kspace."""
        if images is not None:
            self._load(self.images, images)
        self._replay(False, self.norm_code)
        return self.kspace

    def adjoint(self, kspace=None):
        """images[i] = nufft_adj(kspace[i]) for every item; returns self.images."""
        if kspace is not None:
            self._load(self.kspace, kspace)
        self._replay(True, self.norm_code)
        return self.images

    def normal_adjoint_code(self) -> int:
        """Norm code of the exact adjoint of `forward` (for gradient use)."""
        return _adjoint_code(self.norm_code)

    def _load(self, dst, src):
        if len(src) != len(dst):
            raise ValueError(f"expected {len(dst)} items, got {len(src)}")
        srcs = []
        for d, s in zip(dst, src):
            s = _to_torch(s)
            if s.numel() != d.numel():
                raise ValueError(f"item of {s.numel()} values does not fit a buffer of shape {tuple(d.shape)}")
            srcs.append(s.reshape(d.shape).to(device=self.device, dtype=torch.complex64))
        torch._foreach_copy_(dst, srcs)


def plan_nufft_batch(coords, shapes, nbatch=1, width=4, oversamp=1.25, weights=None, basis=None, norm=None,
                     device="cuda", graph=True) -> NUFFTBatch:
    """Plan an uneven batch of NUFFT problems (see NUFFTBatch)."""
    return NUFFTBatch(coords, shapes, nbatch, width, oversamp, weights, basis, norm, device, graph)
