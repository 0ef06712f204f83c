"""pf_nufft_noalloc: the operator without a workspace argument or any
allocation (a fixed per-GPU device array as workspace), called through the C
ABI and checked bit for bit against pf_nufft / pf_nufft_adj with a
caller-allocated workspace, on two calls in a row, the second smaller (so
stale contents of the shared array from the first call must not leak)."""

import ctypes

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from pyfourier_amd import _lib, demo
from pyfourier_amd._nufft import _Call

pytestmark = pytest.mark.gpu


def _ws_call(call, op, src, C, norm_code, weight):
    """The existing entry point with a caller-allocated workspace."""
    L = _lib.lib()
    desc = call.desc(C, (None, None))
    n = ctypes.c_size_t()
    _lib.check(L.pf_workspace_size(call.native.handle, ctypes.byref(desc), op, ctypes.byref(n)))
    ws = torch.empty(max(1, n.value), dtype=torch.uint8, device=call.device)
    it = call.interp
    if op == 0:
        out = torch.empty((C, it.nframes, it.npts), dtype=torch.complex64, device=call.device)
        fn = L.pf_nufft
    else:
        out = torch.empty((C, call.nimg) + tuple(it.shape), dtype=torch.complex64, device=call.device)
        fn = L.pf_nufft_adj
    _lib.check(fn(call.native.handle, ctypes.byref(desc), _lib.ptr(src), _lib.ptr(out), _lib.ptr(weight), norm_code,
                  _lib.ptr(ws), n.value, _lib.stream_of(call.device)))
    return out, n.value


def _noalloc_call(call, op, src, C, norm_code, weight):
    L = _lib.lib()
    desc = call.desc(C, (None, None))
    it = call.interp
    if op == 0:
        out = torch.full((C, it.nframes, it.npts), float("nan"), dtype=torch.complex64, device=call.device)
    else:
        out = torch.full((C, call.nimg) + tuple(it.shape), float("nan"), dtype=torch.complex64, device=call.device)
    rc = L.pf_nufft_noalloc(call.native.handle, ctypes.byref(desc), op, _lib.ptr(src), _lib.ptr(out),
                            _lib.ptr(weight), norm_code, _lib.stream_of(call.device))
    return rc, out


def _case(dev, ndim, N, C, nsp, ns, seed):
    rng = np.random.default_rng(seed)
    coord = demo.golden_angle_radial_2d(nsp, ns) if ndim == 2 else demo.golden_means_radial_3d(nsp, ns)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0).to(dev)
    call = _Call(plan, dev, None, None, None)
    w = torch.from_numpy(rng.uniform(0.5, 2.0, coord.shape[0] * coord.shape[1]).astype(np.float32)).to(dev)
    x = torch.from_numpy(demo.complex_gaussian(rng, (C, 1) + (N,) * ndim)).to(dev)
    y = torch.from_numpy(demo.complex_gaussian(rng, (C, 1, coord.shape[0] * coord.shape[1]))).to(dev)
    return call, w, x, y


def test_noalloc_matches_workspace_entry_points(cuda_device):
    # first call larger than the second: 2-D 128^2 x 8 coils, then 64^2 x 3 coils
    cases = [(2, 128, 8, 96, 256, 0), (2, 64, 3, 40, 128, 1)]
    for ndim, N, C, nsp, ns, seed in cases:
        print(f"noalloc: {ndim}-D N={N} C={C} {nsp}x{ns}", flush=True)
        call, w, x, y = _case(cuda_device, ndim, N, C, nsp, ns, seed)
        for norm_code in (_lib.PF_NORM_NONE, _lib.PF_NORM_ORTHO):
            for op, src in ((0, x), (1, y)):
                ref, need = _ws_call(call, op, src, C, norm_code, w)
                assert need <= _lib.PF_NOALLOC_SCRATCH_BYTES
                rc, got = _noalloc_call(call, op, src, C, norm_code, w)
                assert rc == 0, _lib.lib().pf_last_error().decode()
                torch.cuda.synchronize()
                assert torch.equal(got.view(torch.float32), ref.view(torch.float32)), (N, C, norm_code, op)


def test_noalloc_3d_then_smaller_2d(cuda_device):
    for ndim, N, C, nsp, ns, seed in [(3, 32, 4, 256, 64, 2), (2, 48, 2, 24, 96, 3)]:
        print(f"noalloc: {ndim}-D N={N} C={C} {nsp}x{ns}", flush=True)
        call, w, x, y = _case(cuda_device, ndim, N, C, nsp, ns, seed)
        for op, src in ((0, x), (1, y)):
            ref, _ = _ws_call(call, op, src, C, _lib.PF_NORM_NONE, None)
            rc, got = _noalloc_call(call, op, src, C, _lib.PF_NORM_NONE, None)
            assert rc == 0, _lib.lib().pf_last_error().decode()
            torch.cuda.synchronize()
            assert torch.equal(got.view(torch.float32), ref.view(torch.float32)), (ndim, op)


def test_noalloc_rejects_calls_beyond_its_scratch(cuda_device):
    # 3-D 128^3 x 2 coils at sigma 2: the grid alone is 256^3 x 2 x 8 B = 268 MB
    coord = demo.golden_means_radial_3d(64, 64)
    plan = pf.plan_nufft(torch.from_numpy(coord), 128, width=4, oversamp=2.0).to(cuda_device)
    call = _Call(plan, cuda_device, None, None, None)
    x = torch.zeros((2, 1, 128, 128, 128), dtype=torch.complex64, device=cuda_device)
    out = torch.zeros((2, 1, 64 * 64), dtype=torch.complex64, device=cuda_device)
    desc = call.desc(2, (None, None))
    L = _lib.lib()
    rc = L.pf_nufft_noalloc(call.native.handle, ctypes.byref(desc), 0, _lib.ptr(x), _lib.ptr(out), None,
                            _lib.PF_NORM_NONE, _lib.stream_of(cuda_device))
    assert rc == 5 and "workspace" in L.pf_last_error().decode()
    rc = L.pf_nufft_noalloc(call.native.handle, ctypes.byref(desc), 2, _lib.ptr(x), _lib.ptr(out), None,
                            _lib.PF_NORM_NONE, _lib.stream_of(cuda_device))
    assert rc == 1
