"""Plan autotuner: picks the sort/spread tile and the spread work-item size.

The tile sets how samples are binned (the gather's memory locality) and
which register-resident spread instance runs (visits per sample versus
accumulator rows per visit); the work-item size sets where heavily sampled
tiles (the k-space centre of radial and spiral trajectories) are split into
chunks that a reduce pass sums.  The best choice depends on the trajectory,
the kernel width and the batch, so autotune_plan times nufft + nufft_adj
for every candidate on the caller's trajectory and keeps the fastest.
Results are cached per (geometry, batch, device) for the process.

No counterpart in the reference (its Numba kernels have no tiling).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._plan import plan_nufft, resolve_device

_CACHE: dict = {}


def spread_tiles(ndim: int) -> list:
    """Tiles (z, y, x) that have a register-resident spread for ndim."""
    L = _lib.lib()
    n = ctypes.c_int()
    _lib.check(L.pf_spread_tiles(int(ndim), None, 0, ctypes.byref(n)))
    buf = (ctypes.c_int * (3 * max(1, n.value)))()
    _lib.check(L.pf_spread_tiles(int(ndim), buf, n.value, ctypes.byref(n)))
    return [tuple(buf[3 * i + 3 - ndim: 3 * i + 3]) for i in range(n.value)]


def autotune_plan(coord, shape, width=4, oversamp=1.25, nbatch=1, basis=None, device="cuda", chunks=(0, 1024),
                  reps=3, return_table=False, **plan_kw):
    """plan_nufft(coord, shape, width, oversamp, **plan_kw) with the tile and
    work-item size that make nufft + nufft_adj fastest on this trajectory for
    a batch of `nbatch` (and the subspace `basis`, if given)."""
    from ._nufft import nufft, nufft_adj

    dev = resolve_device(device)
    ndim = int(torch.as_tensor(coord).shape[-1])
    cands = [(t, c) for t in spread_tiles(ndim) for c in chunks]
    key = None
    table = []
    best = None
    for tile, chunk in cands:
        plan = plan_nufft(coord, shape, width, oversamp, tile=tile, chunk=chunk, **plan_kw).to(dev)
        it = plan.interpolator
        if key is None:
            key = (ndim, it.shape, it.width, it.oversamp, it.nframes, it.npts, int(nbatch),
                   None if basis is None else tuple(torch.as_tensor(basis).shape), dev.index)
            if key in _CACHE:
                tile, chunk = _CACHE[key]
                plan = plan_nufft(coord, shape, width, oversamp, tile=tile, chunk=chunk, **plan_kw).to(dev)
                return (plan, []) if return_table else plan
        nimg = torch.as_tensor(basis).shape[1] if basis is not None else it.nframes
        lead = (nbatch, nimg) if (basis is not None or it.nframes > 1) else (nbatch,)
        x = torch.randn(lead + tuple(it.shape), dtype=torch.complex64, device=dev)
        b = None if basis is None else torch.as_tensor(basis).to(dev)
        y = nufft(x, nufft_plan=plan, basis=b, device=dev)
        nufft_adj(y, nufft_plan=plan, basis=b, device=dev)
        st = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            y = nufft(x, nufft_plan=plan, basis=b, device=dev)
            nufft_adj(y, nufft_plan=plan, basis=b, device=dev)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        table.append({"tile": tile, "chunk": chunk, "ms": round(ms, 4)})
        if best is None or ms < best[0]:
            best = (ms, tile, chunk, plan)
        del x, y
    _CACHE[key] = (best[1], best[2])
    return (best[3], table) if return_table else best[3]
