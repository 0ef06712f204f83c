"""Every register-spread tile instance and the heavy-tile chunking give the
restatement's operator; the autotuned plan is one of them."""

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import restate as R
from pyfourier_amd import demo

pytestmark = pytest.mark.gpu

PARITY = 1e-5


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(ndim):
    if ndim == 3:
        return demo.golden_means_radial_3d(160, 40), (20, 24, 28)
    return demo.golden_angle_radial_2d(48, 96), (40, 48)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("chunk", [0, 16])
@pytest.mark.parametrize("C", [1, 8, 12, 16, 20, 64])
def test_all_spread_tiles_parity(cuda_device, ndim, chunk, C):
    # the register spread runs with 8, 16 or 32 values per wave (1, 8, 12
    # and 16 take the small-batch instances); batches of whole 8-, 16- or
    # 32-value chunks (8, 16, 64: two chunks per tile) take the
    # software-pipelined form
    rng = np.random.default_rng(ndim)
    coord, shape = _case(ndim)
    P = R.Plan(coord, shape, width=5, oversamp=1.5)
    x = demo.complex_gaussian(rng, (C,) + shape)
    y = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    yr = R.nufft(x[:, None], P)
    xr = R.nufft_adj(y.reshape(C, 1, -1), P)
    for tile in pf.spread_tiles(ndim):
        plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=5, oversamp=1.5, tile=tile, chunk=chunk)
        plan.to(cuda_device)
        assert tuple(plan.interpolator.native(cuda_device).info.tile)[3 - ndim:] == tuple(tile)
        yg = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
        xg = pf.nufft_adj(torch.from_numpy(y).to(cuda_device), nufft_plan=plan, device=cuda_device)
        assert rel(yg.cpu().numpy().reshape(C, 1, -1), yr) < PARITY, f"forward, tile {tile}"
        assert rel(xg.cpu().numpy().reshape(xr.shape), xr) < PARITY, f"adjoint, tile {tile}"


@pytest.mark.parametrize("ndim,tile", [(2, (32, 32)), (3, (8, 8, 8))])
@pytest.mark.parametrize("C", [1, 8, 20])
def test_lds_spread_large_tiles(cuda_device, ndim, tile, C):
    """Tiles larger than every register instance take the LDS-staged spread."""
    rng = np.random.default_rng(10 + ndim + C)
    coord, shape = _case(ndim)
    P = R.Plan(coord, shape, width=5, oversamp=1.5)
    y = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    xr = R.nufft_adj(y.reshape(C, 1, -1), P)
    for chunk in (0, 16):
        plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=5, oversamp=1.5, tile=tile, chunk=chunk)
        plan.to(cuda_device)
        assert tuple(plan.interpolator.native(cuda_device).info.tile)[3 - ndim:] == tuple(tile)
        xg = pf.nufft_adj(torch.from_numpy(y).to(cuda_device), nufft_plan=plan, device=cuda_device)
        assert rel(xg.cpu().numpy().reshape(xr.shape), xr) < PARITY, f"adjoint, tile {tile}, chunk {chunk}"


def test_autotune_plan(cuda_device):
    rng = np.random.default_rng(9)
    coord, shape = _case(3)
    plan, table = pf.autotune_plan(torch.from_numpy(coord), shape, width=4, oversamp=2.0, nbatch=16,
                                   device=cuda_device, return_table=True)
    assert len(table) == len(pf.spread_tiles(3)) * 2
    assert plan.interpolator.tile in pf.spread_tiles(3)
    best = min(table, key=lambda r: r["ms"])
    assert (best["tile"], best["chunk"]) == (plan.interpolator.tile, plan.interpolator.chunk)
    x = demo.complex_gaussian(rng, (16,) + shape)
    P = R.Plan(coord, shape, width=4, oversamp=2.0)
    yg = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(yg.cpu().numpy().reshape(16, 1, -1), R.nufft(x[:, None], P)) < PARITY
