"""A fresh plan's very first nufft / nufft_adj captured into a HIP graph.

Plan creation builds everything an operator call needs (tap records, spread
lists, the fused FFT's twiddle tables), so the calls only enqueue kernels:
nothing allocates device memory or synchronises on the call path, and the
first call of a new plan can be recorded.  Each replay is checked, with eager
work and synchronisations between the replays, against the eager result of
the same call (bit for bit: every kernel is deterministic).
"""

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from pyfourier_amd import demo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ndim", [2, 3])
def test_first_call_captured(cuda_device, ndim):
    print(f"graph capture of the first call, {ndim}-D", flush=True)
    rng = np.random.default_rng(20 + ndim)
    if ndim == 2:
        N, C = 64, 8
        coord = demo.golden_angle_radial_2d(48, 128)
    else:
        N, C = 24, 4
        coord = demo.golden_means_radial_3d(256, 48)
    w = torch.from_numpy(rng.uniform(0.5, 2.0, coord.shape[:-1]).astype(np.float32)).to(cuda_device)
    x = torch.from_numpy(demo.complex_gaussian(rng, (C,) + (N,) * ndim)).to(cuda_device)
    y_in = torch.from_numpy(demo.complex_gaussian(rng, (C,) + coord.shape[:-1])).to(cuda_device)

    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0)
    plan.to(cuda_device)
    torch.cuda.synchronize()

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        y_g = pf.nufft(x, nufft_plan=plan, weight=w, norm="ortho", device=cuda_device)
        x_g = pf.nufft_adj(y_in, nufft_plan=plan, weight=w, norm="ortho", device=cuda_device)

    y_e = pf.nufft(x, nufft_plan=plan, weight=w, norm="ortho", device=cuda_device)
    x_e = pf.nufft_adj(y_in, nufft_plan=plan, weight=w, norm="ortho", device=cuda_device)
    torch.cuda.synchronize()
    for r in range(3):
        y_g.zero_()
        x_g.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y_g.view(torch.float32), y_e.view(torch.float32)), f"replay {r}: forward differs"
        assert torch.equal(x_g.view(torch.float32), x_e.view(torch.float32)), f"replay {r}: adjoint differs"
        # eager work on the same plan between replays
        pf.nufft_adj(y_e, nufft_plan=plan, device=cuda_device)
