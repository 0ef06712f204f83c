"""GPU tests of the graph-replayed uneven batch (plan_nufft_batch).

Every replay must equal the per-item drop-in calls on the same plans (the
same kernels in the same order: bit-identical), replays are repeated with
eager work and synchronisation in between, and every item is checked against
the CPU restatement, forward and adjoint.
"""

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


def _items(rng, n):
    shapes, coords = [], []
    for i in range(n):
        N = int(rng.choice([32, 48, 64]))
        nsp = int(rng.integers(8, 41))
        shapes.append(N)
        coords.append(demo.golden_angle_radial_2d(nsp, 64))
    return coords, shapes


@pytest.mark.parametrize("norm,group", [(None, True), ("ortho", True), (None, False)])
def test_batch_replays_equal_per_item_calls(cuda_device, norm, group):
    """Grouped items run as frames of one ragged plan: per frame the same
    sample order, tiles and (with a fixed spread chunk size; the automatic
    one depends on the whole list) chunks as the item's own plan, so the
    results are bit-identical to the per-item drop-in calls."""
    rng = np.random.default_rng(0)
    C = 3
    coords, shapes = _items(rng, 7)
    weights = [demo.spiral_density(c) for c in coords]
    batch = pf.plan_nufft_batch(coords, shapes, nbatch=C, width=4, oversamp=2.0, weights=weights, norm=norm,
                                device=cuda_device, group=group, chunk=1024)
    assert (len(batch.groups) > 0) == group
    imgs = [torch.from_numpy(demo.complex_gaussian(rng, (C, 1, N, N))).to(cuda_device) for N in shapes]
    ksps = [torch.from_numpy(demo.complex_gaussian(rng, (C, 1, c.shape[0] * c.shape[1]))).to(cuda_device)
            for c in coords]
    ref_y, ref_x = [], []
    for plan, x, y, w in zip(batch.plans, imgs, ksps, weights):
        wt = torch.from_numpy(w).to(cuda_device)
        ref_y.append(pf.nufft(x[:, 0], nufft_plan=plan, weight=wt, norm=norm, device=cuda_device))
        ref_x.append(pf.nufft_adj(y.reshape((C,) + w.shape), nufft_plan=plan, weight=wt, norm=norm,
                                  device=cuda_device))
    for rep in range(3):
        out = batch.forward(imgs)
        torch.cuda.synchronize()
        for o, r in zip(out, ref_y):
            assert torch.equal(o.reshape(r.shape), r), f"forward replay {rep}"
        # the k-space views are the adjoint's input buffers too: keep copies
        fwd = [o.clone() for o in out]
        # eager work between replays
        junk = torch.randn(1 << 20, device=cuda_device).sum()
        torch.cuda.synchronize()
        outa = batch.adjoint(ksps)
        torch.cuda.synchronize()
        for o, r in zip(outa, ref_x):
            assert torch.equal(o.reshape(r.shape), r), f"adjoint replay {rep}"
        del junk
    # every item against the restatement, forward and adjoint (last replay)
    for i in range(len(coords)):
        P = R.Plan(coords[i], shapes[i], width=4, oversamp=2.0)
        yr = R.nufft(imgs[i].cpu().numpy(), P, weight=weights[i], norm=norm)
        assert rel(fwd[i].cpu().numpy(), yr) < PARITY, f"item {i} forward"
        xr = R.nufft_adj(ksps[i].cpu().numpy(), P, weight=weights[i], norm=norm)
        assert rel(outa[i].cpu().numpy(), xr) < PARITY, f"item {i} adjoint"


def test_batch_input_buffers_in_place(cuda_device):
    """Filling the static buffers in place and replaying gives the same result
    as passing the inputs."""
    rng = np.random.default_rng(1)
    coords, shapes = _items(rng, 4)
    batch = pf.plan_nufft_batch(coords, shapes, nbatch=2, device=cuda_device)
    imgs = [torch.from_numpy(demo.complex_gaussian(rng, (2, 1, N, N))).to(cuda_device) for N in shapes]
    first = [y.clone() for y in batch.forward(imgs)]
    for buf, x in zip(batch.images, imgs):
        buf.copy_(x * 2)
    second = batch.forward()
    torch.cuda.synchronize()
    for a, b in zip(first, second):
        assert torch.allclose(b, 2 * a, rtol=1e-5, atol=1e-5)
    with pytest.raises(ValueError):
        batch.forward(imgs[:-1])


def test_batch_autotuned_groups(cuda_device):
    """autotune=True times every register tile and work-item size on each
    group's own ragged plan and keeps the fastest; the tuned batch still
    matches the restatement item by item, forward and adjoint."""
    rng = np.random.default_rng(3)
    C = 2
    coords, shapes = _items(rng, 5)
    batch = pf.plan_nufft_batch(coords, shapes, nbatch=C, width=4, oversamp=2.0, device=cuda_device, autotune=True)
    ntiles = len(pf.spread_tiles(2))
    for g in batch.groups:
        assert len(g.tune_table) == 2 * ntiles
        best = min(g.tune_table, key=lambda r: r["ms"])
        assert (tuple(best["tile"]), best["chunk"]) == (tuple(g.tile), g.chunk)
    imgs = [torch.from_numpy(demo.complex_gaussian(rng, (C, 1, N, N))).to(cuda_device) for N in shapes]
    ksps = [torch.from_numpy(demo.complex_gaussian(rng, (C, 1, c.shape[0] * c.shape[1]))).to(cuda_device)
            for c in coords]
    fwd = [o.clone() for o in batch.forward(imgs)]
    adj = [o.clone() for o in batch.adjoint(ksps)]
    torch.cuda.synchronize()
    for i in range(len(coords)):
        P = R.Plan(coords[i], shapes[i], width=4, oversamp=2.0)
        assert rel(fwd[i].cpu().numpy(), R.nufft(imgs[i].cpu().numpy(), P)) < PARITY, f"item {i} forward"
        assert rel(adj[i].cpu().numpy(), R.nufft_adj(ksps[i].cpu().numpy(), P)) < PARITY, f"item {i} adjoint"
