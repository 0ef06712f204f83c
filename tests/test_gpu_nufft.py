"""GPU parity of the HIP NUFFT path against the CPU restatement (oracle).

Bar (complex64 engine vs float64 restatement on identical inputs): relative
L2 error <= 1e-5.  Errors against the FP64 NUDFT are kernel-approximation
errors and are checked against the widths' analytic bounds.
"""

import math

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import nudft as U
from oracle import restate as R
from pyfourier_amd import demo

pytestmark = pytest.mark.gpu

PARITY = 1e-5


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _radial2d(nsp=64, ns=256):
    return demo.golden_angle_radial_2d(nsp, ns)


@pytest.mark.parametrize("width,oversamp", [(4, 2.0), (6, 2.0), (4, 1.25), (5, 1.5)])
def test_fwd_adj_2d_parity(cuda_device, width, oversamp):
    rng = np.random.default_rng(0)
    N = 128
    coord = _radial2d()
    x = demo.complex_gaussian(rng, (N, N))
    P = R.Plan(coord, N, width=width, oversamp=oversamp)
    y_ref = R.nufft(x[None, None], P)[0, 0].reshape(coord.shape[:-1])
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), torch.from_numpy(coord), shape=N, width=width,
                 oversamp=oversamp, device=cuda_device)
    assert y.shape == coord.shape[:-1]
    assert rel(y.cpu().numpy(), y_ref) < PARITY
    k = demo.complex_gaussian(rng, coord.shape[:-1])
    x_ref = R.nufft_adj(k.reshape(1, 1, -1), P)[0, 0]
    xa = pf.nufft_adj(torch.from_numpy(k).to(cuda_device), torch.from_numpy(coord), shape=N, width=width,
                      oversamp=oversamp, device=cuda_device)
    assert xa.shape == (N, N)
    assert rel(xa.cpu().numpy(), x_ref) < PARITY


def test_fwd_adj_3d_multicoil_parity(cuda_device):
    rng = np.random.default_rng(1)
    N, C = 32, 5
    coord = demo.golden_means_radial_3d(256, 64)
    x = demo.complex_gaussian(rng, (C, N, N, N))
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    y_ref = R.nufft(x[:, None], P)[:, 0].reshape((C,) + coord.shape[:-1])
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0)
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y.cpu().numpy(), y_ref) < PARITY
    k = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    x_ref = R.nufft_adj(k.reshape(C, 1, -1), P)[:, 0]
    xa = pf.nufft_adj(torch.from_numpy(k).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert xa.shape == (C, N, N, N)
    assert rel(xa.cpu().numpy(), x_ref) < PARITY


@pytest.mark.parametrize("shape,oversamp", [((128, 16, 16), 1.25), ((96, 12, 10), 2.0), ((192, 12, 10), 1.25),
                                           ((256, 12, 10), 1.25), ((192, 12, 10), 2.0), ((64, 12, 10), 2.0),
                                           ((384, 12, 10), 1.25), ((512, 12, 10), 1.25), ((320, 12, 10), 1.25),
                                           ((144, 12, 10), 2.0), ((64, 12, 10), 1.25), ((48, 12, 10), 2.0),
                                           ((96, 12, 10), 1.25), ((160, 12, 10), 1.25), ((180, 12, 10), 2.0),
                                           ((112, 12, 10), 2.0), ((224, 12, 10), 1.25), ((32, 32, 32), 1.25)])
def test_fwd_adj_3d_register_z_pass(cuda_device, shape, oversamp):
    """3-D plans whose z axis (160, 192, 240, 320, 384, 128, 480, 640, 400, 288, 80, 96, 120,
    200, 360, 224, 280 points) runs the
    register column pass, masked in the fused adjoint (the radial sphere
    leaves the (y, x) grid's corner columns empty)."""
    rng = np.random.default_rng(11)
    C = 4
    coord = demo.golden_means_radial_3d(128, 64)
    x = demo.complex_gaussian(rng, (C,) + shape)
    P = R.Plan(coord, shape, width=4, oversamp=oversamp)
    y_ref = R.nufft(x[:, None], P)[:, 0].reshape((C,) + coord.shape[:-1])
    plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=4, oversamp=oversamp)
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y.cpu().numpy(), y_ref) < PARITY
    k = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    x_ref = R.nufft_adj(k.reshape(C, 1, -1), P)[:, 0]
    xa = pf.nufft_adj(torch.from_numpy(k).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(xa.cpu().numpy(), x_ref) < PARITY


@pytest.mark.parametrize("C", [1, 2, 8, 16, 32])
def test_batch_sizes_2d(cuda_device, C):
    rng = np.random.default_rng(2)
    N = 64
    coord = _radial2d(32, 128)
    x = demo.complex_gaussian(rng, (C, N, N))
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    y_ref = R.nufft(x[:, None], P)[:, 0].reshape((C,) + coord.shape[:-1])
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0)
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y.cpu().numpy(), y_ref) < PARITY
    xa = pf.nufft_adj(torch.from_numpy(y_ref.astype(np.complex64)).to(cuda_device), nufft_plan=plan,
                      device=cuda_device)
    x_ref = R.nufft_adj(y_ref.reshape(C, 1, -1), P)[:, 0]
    assert rel(xa.cpu().numpy(), x_ref) < PARITY


def test_adjointness_ortho(cuda_device):
    rng = np.random.default_rng(3)
    N = 96
    coord = _radial2d(48, 192)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0)
    x = torch.from_numpy(demo.complex_gaussian(rng, (3, N, N))).to(cuda_device)
    y = torch.from_numpy(demo.complex_gaussian(rng, (3,) + coord.shape[:-1])).to(cuda_device)
    Ax = pf.nufft(x, nufft_plan=plan, norm="ortho", device=cuda_device)
    Ahy = pf.nufft_adj(y, nufft_plan=plan, norm="ortho", device=cuda_device)
    lhs = torch.vdot(Ax.reshape(-1).cdouble(), y.reshape(-1).cdouble())
    rhs = torch.vdot(x.reshape(-1).cdouble(), Ahy.reshape(-1).cdouble())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_adjoint_deterministic(cuda_device):
    rng = np.random.default_rng(4)
    N = 64
    coord = demo.golden_means_radial_3d(512, 64)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0)
    y = torch.from_numpy(demo.complex_gaussian(rng, (8,) + coord.shape[:-1])).to(cuda_device)
    a = pf.nufft_adj(y, nufft_plan=plan, device=cuda_device)
    for _ in range(3):
        b = pf.nufft_adj(y, nufft_plan=plan, device=cuda_device)
        assert torch.equal(a.view(torch.float32), b.view(torch.float32))


def test_nudft_accuracy_2d(cuda_device):
    rng = np.random.default_rng(5)
    N = 64
    coord = _radial2d(32, 128)
    x = demo.complex_gaussian(rng, (N, N))
    for width, bound in [(4, 2e-3), (6, 5e-5)]:
        plan = pf.plan_nufft(torch.from_numpy(coord), N, width=width, oversamp=2.0)
        y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
        ct = plan.interpolator.coord.reshape(-1, 2).numpy()
        yt = U.nudft(x[None], ct, [N, N])[0]
        assert rel(y.cpu().numpy().ravel(), yt) < bound


@pytest.mark.parametrize("C,nsp,ns", [(4, 36, 64), (4, 37, 63), (3, 36, 64), (2, 1, 1)],
                         ids=["16-byte", "odd-samples", "odd-batch", "one-sample"])
def test_permutation_paths(cuda_device, C, nsp, ns):
    """Sample permutations between user and tile order: 16-byte pieces for
    even batch and sample counts, 8-byte pieces otherwise; with density
    weights on both sides."""
    rng = np.random.default_rng(C * 100 + nsp)
    N = 40
    coord = _radial2d(nsp, ns)
    w = rng.uniform(0.5, 2.0, coord.shape[:-1]).astype(np.float32)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    x = demo.complex_gaussian(rng, (C, N, N))
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0)
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, weight=w, device=cuda_device)
    yr = R.nufft(x[:, None], P, weight=w)[:, 0].reshape((C,) + coord.shape[:-1])
    assert rel(y.cpu().numpy(), yr) < PARITY
    k = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    xa = pf.nufft_adj(torch.from_numpy(k).to(cuda_device), nufft_plan=plan, weight=w, device=cuda_device)
    xr = R.nufft_adj(k.reshape(C, 1, -1), P, weight=w)[:, 0]
    assert rel(xa.cpu().numpy(), xr) < PARITY


def test_units_keyword_on_the_operators(cuda_device):
    """A grid-unit calibration region (radius <= pi, which "auto" would read
    as radians) passed straight to nufft / nufft_adj with units="grid": the
    operator equals the restatement on the unscaled grid coordinates."""
    rng = np.random.default_rng(12)
    N, C = 48, 2
    coord = (rng.uniform(-3.0, 3.0, (4, 96, 2))).astype(np.float32)  # grid units, |k| < pi
    x = demo.complex_gaussian(rng, (C, N, N))
    k = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), torch.from_numpy(coord), width=4, oversamp=2.0,
                 device=cuda_device, units="grid")
    xa = pf.nufft_adj(torch.from_numpy(k).to(cuda_device), torch.from_numpy(coord), shape=N, width=4,
                      oversamp=2.0, device=cuda_device, units="grid")
    P = R.Plan(coord, N, width=4, oversamp=2.0, normalize=False)
    yr = R.nufft(x[:, None], P)[:, 0].reshape((C,) + coord.shape[:-1])
    xr = R.nufft_adj(k.reshape(C, 1, -1), P)[:, 0]
    assert rel(y.cpu().numpy(), yr) < PARITY
    assert rel(xa.cpu().numpy(), xr) < PARITY
    # the default reads the same numbers as radians: a different operator
    y_auto = pf.nufft(torch.from_numpy(x).to(cuda_device), torch.from_numpy(coord), width=4, oversamp=2.0,
                      device=cuda_device)
    assert rel(y_auto.cpu().numpy(), yr) > 0.1
