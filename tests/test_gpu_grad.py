"""GPU parity of the trajectory (coordinate) gradients of nufft / nufft_adj.

The expected values come from the CPU restatement's exact derivative of the
discrete operator (oracle.restate.coord_grad), which tests/test_oracle.py
pins against torch autograd through the exact NUDFT.  Gradients are driven
with an explicit upstream gradient (``backward(gradient=...)``) so that
coord.grad = Re(sum conj(upstream) * d output / d coord).
"""

import math

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import restate as R
from pyfourier_amd import demo

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4
PARITY = 1e-5


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _radian_units(shape_zyx):
    """d(grid units)/d(radians) per coordinate axis (x, y, z)."""
    return np.asarray(shape_zyx[::-1], dtype=np.float64) / (2 * math.pi)


def test_forward_2d_radial_radians(cuda_device):
    rng = np.random.default_rng(0)
    N, C = 48, 3
    coord = demo.golden_angle_radial_2d(24, 96)
    x = demo.complex_gaussian(rng, (C, N, N))
    v = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    ct = torch.from_numpy(coord).requires_grad_()
    xt = torch.from_numpy(x).to(cuda_device).requires_grad_()
    y = pf.nufft(xt, ct, device=cuda_device)
    vt = torch.from_numpy(v).to(cuda_device)
    y.backward(gradient=vt)
    P = R.Plan(coord, N, width=4, oversamp=1.25)
    want = R.coord_grad(x[:, None], P, v.reshape(C, 1, -1)) * _radian_units((N, N))
    assert ct.grad.shape == ct.shape and ct.grad.dtype == ct.dtype
    assert rel(ct.grad.numpy().reshape(-1, 2), want) < GRAD_TOL
    # the image gradient is the adjoint of the upstream gradient
    xa = R.nufft_adj(v.reshape(C, 1, -1), P)
    G = float(np.prod(P.os_shape))
    assert rel(xt.grad.cpu().numpy(), xa.reshape(x.shape) * G) < PARITY


def test_adjoint_3d_ortho_weighted(cuda_device):
    rng = np.random.default_rng(1)
    shape = (20, 24, 28)
    C = 2
    coord = demo.golden_means_radial_3d(120, 40)
    w = rng.uniform(0.2, 1.0, coord.shape[:-1]).astype(np.float32)
    y = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    u = demo.complex_gaussian(rng, (C,) + shape)
    ct = torch.from_numpy(coord).requires_grad_()
    yt = torch.from_numpy(y).to(cuda_device).requires_grad_()
    x = pf.nufft_adj(yt, ct, shape=shape, weight=torch.from_numpy(w), norm="ortho", width=5, oversamp=1.5,
                     device=cuda_device)
    x.backward(gradient=torch.from_numpy(u).to(cuda_device))
    P = R.Plan(coord, shape, width=5, oversamp=1.5)
    want = R.coord_grad(u[:, None], P, y.reshape(C, 1, -1), weight=w, norm="ortho", adjoint=True)
    assert rel(ct.grad.numpy().reshape(-1, 3), want * _radian_units(shape)) < GRAD_TOL
    # k-space gradient: the forward operator applied to the upstream gradient
    yk = R.nufft(u[:, None], P, weight=w, norm="ortho")
    assert rel(yt.grad.cpu().numpy().reshape(C, 1, -1), yk) < PARITY


@pytest.mark.parametrize(
    "width,N,C,F",
    [
        (3, 20, 4, 1),
        (6, 24, 6, 1),
        (6, 16, 4, 2),  # two frames: one grid per frame
        (8, 16, 2, 1),
    ],
)
def test_forward_3d_staged_window(cuda_device, width, N, C, F):
    """3-D plans with one width on all axes and an even batch: the
    derivative gather reads the coil-pair-major grid through the LDS-staged
    window (k_gather_dcoord_col<W>), one lane per sample over all coil pairs."""
    print(f"dcoord col: W={width} N={N} C={C} F={F}", flush=True)
    rng = np.random.default_rng(40 + width + C)
    base = demo.golden_means_radial_3d(60 * F, 32)
    coord = np.stack([base[f::F] for f in range(F)]) if F > 1 else base
    x = demo.complex_gaussian(rng, (C,) + ((F,) if F > 1 else ()) + (N, N, N))
    v = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    ct = torch.from_numpy(coord).requires_grad_()
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), ct, width=width, oversamp=2.0, device=cuda_device)
    y.backward(gradient=torch.from_numpy(v).to(cuda_device))
    P = R.Plan(coord, N, width=width, oversamp=2.0)
    xr = x.reshape((C, F, N, N, N))
    want = R.coord_grad(xr, P, v.reshape(C, F, -1)) * _radian_units((N, N, N))
    assert rel(ct.grad.numpy().reshape(-1, 3), want) < GRAD_TOL


def test_subspace_multiframe_cycles(cuda_device):
    rng = np.random.default_rng(2)
    N, F, K, C = 32, 4, 2, 2
    base = demo.golden_angle_radial_2d(4 * F, 64) / (2 * math.pi)  # cycles per FOV
    coord = np.ascontiguousarray(base.reshape(F, 4, 64, 2))
    basis = demo.t2_subspace_basis(rng, F, K)
    x = demo.complex_gaussian(rng, (C, K, N, N))
    v = demo.complex_gaussian(rng, (C, F, 4, 64))
    ct = torch.from_numpy(coord).requires_grad_()
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), ct, basis=torch.from_numpy(basis), device=cuda_device)
    y.backward(gradient=torch.from_numpy(v).to(cuda_device))
    P = R.Plan(coord, N, width=4, oversamp=1.25)
    want = R.coord_grad(x, P, v.reshape(C, F, -1), basis=basis) * float(N)
    assert rel(ct.grad.numpy().reshape(-1, 2), want) < GRAD_TOL


def test_b0_segments_forward_and_adjoint(cuda_device):
    rng = np.random.default_rng(3)
    N, C, L = 32, 2, 5
    coord = demo.variable_density_spiral_2d(6, 256, N)
    zmap = demo.smooth_field_map(rng, (N, N), 80.0)
    t = np.linspace(0, 8e-3, 256).astype(np.float32)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0, zmap=torch.from_numpy(zmap),
                         T=torch.from_numpy(t), L=L, nbins=40, L_batch_size=2)
    bt = plan.zmap_t_kernel.cpu().numpy().astype(np.complex128)
    cs = plan.zmap_s_kernel.cpu().numpy().astype(np.complex128)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    x = demo.complex_gaussian(rng, (C, N, N))
    v = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    # forward, with the plan given and coord marked for differentiation
    ct = torch.from_numpy(coord).requires_grad_()
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), ct, nufft_plan=plan, device=cuda_device)
    y.backward(gradient=torch.from_numpy(v).to(cuda_device))
    want = R.coord_grad(x[:, None], P, v.reshape(C, 1, -1), b0=(bt, cs)) * _radian_units((N, N))
    assert rel(ct.grad.numpy().reshape(-1, 2), want) < GRAD_TOL
    # adjoint
    ct = torch.from_numpy(coord).requires_grad_()
    xa = pf.nufft_adj(torch.from_numpy(v).to(cuda_device), ct, nufft_plan=plan, device=cuda_device)
    xa.backward(gradient=torch.from_numpy(x).to(cuda_device))
    want = R.coord_grad(x[:, None], P, v.reshape(C, 1, -1), b0=(bt, cs), adjoint=True) * _radian_units((N, N))
    assert rel(ct.grad.numpy().reshape(-1, 2), want) < GRAD_TOL


def test_grid_unit_coordinates_and_gradient_descent_step(cuda_device):
    """Coordinates already in grid units get unit scale 1; a small step along
    -grad lowers a data-consistency loss (first-order check of the sign)."""
    rng = np.random.default_rng(4)
    N = 24
    coord = (rng.uniform(-0.5, 0.5, (400, 2)) * N).astype(np.float32)
    coord[0] = [-N / 2, -N / 2]  # radius > pi: grid units
    x = demo.complex_gaussian(rng, (N, N))
    target = pf.nufft(torch.from_numpy(x).to(cuda_device), torch.from_numpy(coord + 0.05), device=cuda_device)

    def loss(c):
        y = pf.nufft(torch.from_numpy(x).to(cuda_device), c, device=cuda_device)
        return (y - target).abs().pow(2).sum()

    ct = torch.from_numpy(coord).requires_grad_()
    L0 = loss(ct)
    L0.backward()
    P = R.Plan(coord, N, width=4, oversamp=1.25)
    y0 = R.nufft(x[None, None].astype(np.complex128), P)
    h = 2 * (y0 - target.cpu().numpy().reshape(y0.shape))  # dL/dy* convention of torch: 2 (y - t)
    want = R.coord_grad(x[None, None], P, h)
    assert rel(ct.grad.numpy(), want) < GRAD_TOL
    step = 1e-3 / float(ct.grad.abs().max())
    with torch.no_grad():
        L1 = loss(ct - step * ct.grad)
    assert float(L1) < float(L0)


def test_coord_grad_vs_nudft_gradient(cuda_device):
    """Accurate kernel (W=6, sigma=2): the trajectory gradient of the engine
    against torch autograd through the exact FP64 NUDFT."""
    rng = np.random.default_rng(5)
    N, C, M = 24, 2, 300
    coord = (rng.uniform(-0.5, 0.5, (M, 2)) * N).astype(np.float32)
    coord[0] = [-N / 2, -N / 2]
    x = demo.complex_gaussian(rng, (C, N, N))
    v = demo.complex_gaussian(rng, (C, M))
    ct = torch.from_numpy(coord).requires_grad_()
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), ct, width=6, oversamp=2.0, device=cuda_device)
    y.backward(gradient=torch.from_numpy(v).to(cuda_device).reshape(y.shape))
    # exact NUDFT y_j = sum_r x_r exp(-2 pi i k_j . r / N), differentiated by torch in FP64
    kt = torch.from_numpy(coord.astype(np.float64)).requires_grad_()
    r = torch.arange(N, dtype=torch.float64) - N // 2
    ph = kt[:, 0, None, None] * r[None, None, :] / N + kt[:, 1, None, None] * r[None, :, None] / N
    E = torch.exp(-2j * math.pi * ph)
    ye = (E[None] * torch.from_numpy(x.astype(np.complex128))[:, None]).flatten(2).sum(-1)
    ye.backward(gradient=torch.from_numpy(v.astype(np.complex128)))
    assert rel(ct.grad.numpy(), kt.grad.numpy()) < 1e-3
