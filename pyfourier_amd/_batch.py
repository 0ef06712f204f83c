"""Uneven batches of independent NUFFT problems in a few launches.

A reconstruction or training loop over many slices, each with its own
trajectory and sample count, pays per slice both the host cost of the
drop-in API and, worse, kernels too small to fill the GPU.  NUFFTBatch
plans every item once and groups the items that share an image geometry
(matrix, kernel width, oversampling): a group runs as ONE ragged plan whose
frames are the items (pf_plan_create_ragged), so each kernel of the
pipeline covers every sample of the group in one launch.  Items that cannot
join a group (several frames, or a subspace basis) run as their own calls.
All launches of a direction are recorded once into a torch.cuda.CUDAGraph
(a hipGraph on ROCm) and replayed.

No counterpart in the reference (which loops in Python); the operator of
each item is exactly the one nufft / nufft_adj apply with that item's plan.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._nufft import _Call, _adjoint_code, _norm_code, _to_torch
from ._plan import NativePlan, plan_nufft, resolve_device


class _Group:
    """Items of one geometry as the frames of one ragged native plan.

    autotune: build the ragged plan for every register-spread tile and
    spread work-item size (as autotune_plan does for one plan), time this
    group's forward + adjoint with its batch on each, and keep the fastest;
    `tune_table` then lists the candidates' times."""

    def __init__(self, plans, weights, nbatch, device, tile, chunk, autotune=False, chunks=(0, 1024), reps=3):
        it0 = plans[0].interpolator
        d = it0.ndim
        coords = [p.interpolator.coord.reshape(-1, d) for p in plans]
        counts = [int(c.shape[0]) for c in coords]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.device = device
        cat = torch.cat(coords)

        def build(t, c):
            return NativePlan(cat, len(plans), 0, it0.shape, it0.width, it0.oversamp, device, t, c,
                              frame_offsets=off.tolist())

        total = int(off[-1])
        self.nbatch = nbatch
        self.tune_table = []
        if autotune:
            from ._tune import spread_tiles

            self.images = torch.randn((nbatch, len(plans)) + tuple(it0.shape), dtype=torch.complex64, device=device)
            self.kspace = torch.empty((nbatch, total), dtype=torch.complex64, device=device)
            self.weight = None
            best = None
            for t in spread_tiles(d):
                for c in chunks:
                    self.native = build(t, c)
                    ms = self._time(reps)
                    self.tune_table.append({"tile": t, "chunk": c, "ms": round(ms, 4)})
                    if best is None or ms < best[0]:
                        best = (ms, t, c)
            tile, chunk = best[1], best[2]
        self.tile, self.chunk = tile, chunk
        self.native = build(tile, chunk)
        self.images = torch.zeros((nbatch, len(plans)) + tuple(it0.shape), dtype=torch.complex64, device=device)
        self.kspace = torch.zeros((nbatch, total), dtype=torch.complex64, device=device)
        self.weight = None
        if any(w is not None for w in weights):
            parts = [torch.ones(n, dtype=torch.float32, device=device) if w is None else w
                     for w, n in zip(weights, counts)]
            self.weight = torch.cat(parts).contiguous()
        self.image_views = [self.images[:, k:k + 1] for k in range(len(plans))]
        self.kspace_views = [self.kspace[:, int(off[k]):int(off[k + 1])].unsqueeze(1) for k in range(len(plans))]

    def _time(self, reps):
        """ms per forward + adjoint of the current native plan."""
        self._run(False, 0)
        self._run(True, _adjoint_code(0))
        st = torch.cuda.current_stream(self.device)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            self._run(False, 0)
            self._run(True, _adjoint_code(0))
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def _run(self, adjoint, code):
        L = _lib.lib()
        desc = _lib.BatchDesc()
        desc.batch, desc.nseg, desc.ncoeff, desc.nt = self.nbatch, 1, 0, 1
        nbytes = ctypes.c_size_t()
        _lib.check(L.pf_workspace_size(self.native.handle, ctypes.byref(desc), int(adjoint), ctypes.byref(nbytes)))
        ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.device)
        fn = L.pf_nufft_adj if adjoint else L.pf_nufft
        src, dst = (self.kspace, self.images) if adjoint else (self.images, self.kspace)
        with torch.cuda.device(self.device):
            _lib.check(fn(self.native.handle, ctypes.byref(desc), _lib.ptr(src), _lib.ptr(dst),
                          _lib.ptr(self.weight), code, _lib.ptr(ws), nbytes.value, _lib.stream_of(self.device)))


class NUFFTBatch:
    """Forward / adjoint NUFFT of an uneven list of problems.

    coords[i]: trajectory of item i, (ncontrasts, nviews, nsamples, ndim) or
    fewer leading axes, in any unit; shapes[i]: its image matrix.  Every
    item carries `nbatch` batch entries (coils).  weights[i] (optional) is a
    real density weight broadcastable to item i's k-space; basis (optional,
    shared) is a (ncontrasts, ncoeff) subspace basis.

    ``images[i]`` (nbatch, nimg_i, *shape_i) and ``kspace[i]`` (nbatch,
    ncontrasts_i, npts_i) are (views of) the static buffers the launches
    read and write: fill them in place, or pass lists to forward/adjoint,
    which copy with one multi-tensor copy.  A call overwrites the previous
    results.  group=False runs every item as its own calls (one graph).
    """

    def __init__(self, coords, shapes, nbatch=1, width=4, oversamp=1.25, weights=None, basis=None, norm=None,
                 device="cuda", graph=True, group=True, tile=None, chunk=0, units="auto", autotune=False):
        if len(coords) != len(shapes):
            raise ValueError("coords and shapes must have the same length")
        if len(coords) == 0:
            raise ValueError("empty batch")
        self.device = resolve_device(device)
        self.norm_code = _norm_code(norm)
        self.use_graph = bool(graph)
        self.plans = [plan_nufft(_to_torch(c).detach().to(torch.float32), sh, width, oversamp, tile=tile, chunk=chunk,
                                 units=units)
                      for c, sh in zip(coords, shapes)]
        self.images = [None] * len(coords)
        self.kspace = [None] * len(coords)
        self.calls = {}
        self.groups = []
        wts = [None] * len(coords)
        groups = {}
        for i, plan in enumerate(self.plans):
            it = plan.interpolator
            w = None if weights is None else weights[i]
            if w is not None:
                w = _to_torch(w).to(self.device)
                if torch.is_complex(w):
                    raise ValueError("NUFFTBatch takes real density weights")
                w = torch.broadcast_to(w.to(torch.float32), (it.nframes,) + tuple(it.dshape)).reshape(-1)
                wts[i] = w.contiguous()
            if group and basis is None and it.nframes == 1:
                groups.setdefault((it.ndim, it.shape, it.width, it.oversamp), []).append(i)
            else:
                plan.to(self.device)
                call = _Call(plan, self.device, basis, None, norm)
                call.weight_flat = wts[i]
                self.calls[i] = call
                self.images[i] = torch.zeros((nbatch, call.nimg) + tuple(it.shape), dtype=torch.complex64,
                                             device=self.device)
                self.kspace[i] = torch.zeros((nbatch, it.nframes, it.npts), dtype=torch.complex64,
                                             device=self.device)
        for idx in groups.values():
            g = _Group([self.plans[i] for i in idx], [wts[i] for i in idx], nbatch, self.device, tile, chunk,
                       autotune=autotune and tile is None)
            self.groups.append(g)
            for k, i in enumerate(idx):
                self.images[i] = g.image_views[k]
                self.kspace[i] = g.kspace_views[k]
        self._graphs = {}

    def __len__(self):
        return len(self.plans)

    # ---- one pass over all items, eager ----------------------------------
    def _launch(self, adjoint: bool, code: int):
        for g in self.groups:
            g._run(adjoint, code)
        for i, call in self.calls.items():
            if adjoint:
                call.adjoint(self.kspace[i], code, out=self.images[i])
            else:
                call.forward(self.images[i], code, out=self.kspace[i])

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
                     device="cuda", graph=True, group=True, tile=None, chunk=0, units="auto",
                     autotune=False) -> NUFFTBatch:
    """Plan an uneven batch of NUFFT problems (see NUFFTBatch).  autotune=True
    (with tile=None) picks each group's tile and spread work-item size by
    timing the group's own forward + adjoint on every candidate."""
    return NUFFTBatch(coords, shapes, nbatch, width, oversamp, weights, basis, norm, device, graph, group, tile,
                      chunk, units, autotune)
