"""The LDS-staged gather (k_gather_col<W>) against the oracle.

It runs for 3-D plans with one kernel width W (3..8) on all axes, plain
values and an even batch (two coils per lane); the window of
every work item is a 16 x 16 column in (y, x) by a few z planes plus the
kernel halo, wrapped periodically.  Cases: every width, grids smaller than
the staged window (the wrap repeats points), several frames (one grid per
frame), the batch extremes 2 and 64, density weights, and the instance the
plan reports (pf_describe_launch).  Reference: the float64 restatement of
src/pyfourier/_subroutines/_grid/_cart2noncart/_3D.py:13-51 inside the
pipeline of src/pyfourier/_nufft/_nufft.py:218-251, relative L2 <= 1e-5.
Odd batches and unequal widths keep the direct-read gather (covered by the
other GPU tests).
"""

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import restate as R
from pyfourier_amd import demo
from pyfourier_amd._nufft import _Call

pytestmark = pytest.mark.gpu

PARITY = 1e-5


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize(
    "width,N,C,nsp,ns,oversamp",
    [
        (3, 24, 4, 200, 48, 2.0),
        (4, 32, 8, 256, 64, 2.0),
        (5, 20, 6, 160, 40, 1.5),
        (6, 32, 2, 256, 64, 2.0),
        (6, 24, 64, 128, 48, 2.0),
        (7, 16, 4, 128, 32, 2.0),
        (8, 16, 4, 128, 32, 2.0),
        (6, 8, 4, 64, 16, 2.0),  # 16^3 grid: smaller than the 21 x 21 x 11 window
        (4, 40, 2, 200, 80, 1.25),  # grid 50: columns and windows cut by the edge
    ],
)
def test_gather_col_vs_oracle(cuda_device, width, N, C, nsp, ns, oversamp):
    print(f"gather_col: W={width} N={N} C={C} {nsp}x{ns} sigma={oversamp}", flush=True)
    rng = np.random.default_rng(100 + width * 7 + C)
    coord = demo.golden_means_radial_3d(nsp, ns)
    w = rng.uniform(0.5, 2.0, coord.shape[:-1]).astype(np.float32)
    x = demo.complex_gaussian(rng, (C, N, N, N))
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=width, oversamp=oversamp).to(cuda_device)
    what = _Call(plan, cuda_device, None, None, None).describe(C)
    assert f"gather=k_gather_col<{width}>" in what, what
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, weight=torch.from_numpy(w),
                 device=cuda_device)
    P = R.Plan(coord, N, width=width, oversamp=oversamp)
    yr = R.nufft(x[:, None], P, weight=w)[:, 0].reshape((C,) + coord.shape[:-1])
    assert rel(y.cpu().numpy(), yr) < PARITY


def test_gather_col_frames(cuda_device):
    """Several frames: one grid per frame, the frame is part of the window key."""
    rng = np.random.default_rng(5)
    F, N, C = 3, 24, 4
    base = demo.golden_means_radial_3d(300, 48)
    coord = np.stack([base[f::F] for f in range(F)])  # (F, 100, 48, 3)
    x = demo.complex_gaussian(rng, (C, F, N, N, N))
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0).to(cuda_device)
    assert "gather=k_gather_col<6>" in _Call(plan, cuda_device, None, None, None).describe(C)
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    P = R.Plan(coord, N, width=6, oversamp=2.0)
    yr = R.nufft(x, P).reshape((C, F) + coord.shape[1:-1])
    assert rel(y.cpu().numpy(), yr) < PARITY


def test_gather_col_not_used_where_it_does_not_apply(cuda_device):
    coord = demo.golden_means_radial_3d(64, 32)
    plan = pf.plan_nufft(torch.from_numpy(coord), 16, width=4, oversamp=2.0).to(cuda_device)
    call = _Call(plan, cuda_device, None, None, None)
    assert "k_gather_col" not in call.describe(3)  # odd batch
    aniso = pf.plan_nufft(torch.from_numpy(coord), 16, width=(4, 6, 6), oversamp=2.0).to(cuda_device)
    assert "k_gather_col" not in _Call(aniso, cuda_device, None, None, None).describe(4)
