"""GPU tests of the Toeplitz Gram (plan_spgram / sparse_gram).

The Gram must equal the normal operator nufft_adj(w * nufft(x)) of the NUFFT
it replaces up to kernel-approximation error, and the oracle's Toeplitz apply
with the FP64-exact PSF up to the PSF's own NUFFT error.
"""

import math

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import restate as R
from pyfourier_amd import demo

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_gram_vs_exact_and_explicit(cuda_device):
    rng = np.random.default_rng(0)
    N, C = 32, 3
    coord = demo.variable_density_spiral_2d(6, 400, N)
    w = demo.spiral_density(coord)
    G = pf.plan_spgram(torch.from_numpy(coord), N, weight=torch.from_numpy(w), oversamp=2.0, device=cuda_device)
    x = demo.complex_gaussian(rng, (C, N, N))
    gx = pf.sparse_gram(torch.from_numpy(x).to(cuda_device), G, device=cuda_device).cpu().numpy()
    # oracle: exact PSF (FP64 NUDFT) Toeplitz, normal operator of a sigma=2 NUFFT
    P = R.Plan(coord, N, width=6, oversamp=2.0)
    T = R.toeplitz_kernel_exact(P.coord.reshape(-1, 2), (N, N), w.ravel(), ghat=float(np.prod(P.os_shape)))
    ref = np.stack([R.toeplitz_apply(x[c], T, (N, N)) for c in range(C)])
    assert rel(gx, ref) < 1e-4
    # explicit normal operator A^H W A of the engine's own NUFFT (W=6, sigma=2);
    # the weight enters once, as in the reference PSF (weight * delta,
    # src/pyfourier/_sparsegram/_plan.py:202)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0)
    xd = torch.from_numpy(x).to(cuda_device)
    wd = torch.from_numpy(w).to(cuda_device)
    ex = pf.nufft_adj(pf.nufft(xd, nufft_plan=plan, weight=wd, device=cuda_device), nufft_plan=plan,
                      device=cuda_device)
    assert rel(gx, ex.cpu().numpy()) < 1e-4


def test_gram_3d_hermitian(cuda_device):
    rng = np.random.default_rng(1)
    N = 16
    coord = demo.golden_means_radial_3d(200, 32)
    G = pf.plan_spgram(torch.from_numpy(coord), N, oversamp=2.0, device=cuda_device)
    x = torch.from_numpy(demo.complex_gaussian(rng, (N, N, N))).to(cuda_device)
    y = torch.from_numpy(demo.complex_gaussian(rng, (N, N, N))).to(cuda_device)
    gx = pf.sparse_gram(x, G, device=cuda_device)
    gy = pf.sparse_gram(y, G, device=cuda_device)
    a = torch.vdot(y.reshape(-1).cdouble(), gx.reshape(-1).cdouble())
    b = torch.vdot(gy.reshape(-1).cdouble(), x.reshape(-1).cdouble())
    assert abs(a - b) / abs(a) < 1e-5
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0)
    ex = pf.nufft_adj(pf.nufft(x, nufft_plan=plan, device=cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(gx.cpu().numpy(), ex.cpu().numpy()) < 2e-4


def test_gram_subspace(cuda_device):
    rng = np.random.default_rng(2)
    F, K, N, C = 12, 3, 24, 2
    base = demo.golden_angle_radial_2d(2, 64)
    coord = np.stack([base @ np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]]).T
                      for a in 0.61 * np.arange(F)]).astype(np.float32)
    basis = demo.t2_subspace_basis(rng, F, K)
    G = pf.plan_spgram(torch.from_numpy(coord), N, basis=torch.from_numpy(basis), oversamp=2.0,
                       device=cuda_device)
    x = torch.from_numpy(demo.complex_gaussian(rng, (C, K, N, N))).to(cuda_device)
    gx = pf.sparse_gram(x, G, device=cuda_device)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0)
    b = torch.from_numpy(basis)
    ex = pf.nufft_adj(pf.nufft(x, nufft_plan=plan, basis=b, device=cuda_device), nufft_plan=plan, basis=b,
                      device=cuda_device)
    assert rel(gx.cpu().numpy(), ex.cpu().numpy()) < 2e-4


def test_gram_b0(cuda_device):
    rng = np.random.default_rng(3)
    N, C, L = 32, 2, 4
    coord = demo.variable_density_spiral_2d(4, 512, N)
    zmap = demo.smooth_field_map(rng, (N, N), 80.0)
    t = np.linspace(0, 8e-3, 512).astype(np.float32)
    G = pf.plan_spgram(torch.from_numpy(coord), N, zmap=torch.from_numpy(zmap), T=torch.from_numpy(t), L=L,
                       nbins=40, oversamp=2.0, device=cuda_device)
    x = torch.from_numpy(demo.complex_gaussian(rng, (C, N, N))).to(cuda_device)
    gx = pf.sparse_gram(x, G, device=cuda_device)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0, zmap=torch.from_numpy(zmap),
                         T=torch.from_numpy(t), L=L, nbins=40)
    ex = pf.nufft_adj(pf.nufft(x, nufft_plan=plan, device=cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(gx.cpu().numpy(), ex.cpu().numpy()) < 2e-4


def test_gram_autograd_is_self_adjoint(cuda_device):
    rng = np.random.default_rng(4)
    N = 24
    coord = demo.golden_angle_radial_2d(30, 64)
    G = pf.plan_spgram(torch.from_numpy(coord), N, oversamp=2.0, device=cuda_device)
    x = torch.from_numpy(demo.complex_gaussian(rng, (N, N))).to(cuda_device).requires_grad_(True)
    y = pf.sparse_gram(x, G, device=cuda_device)
    loss = 0.5 * (y.abs() ** 2).sum()
    loss.backward()
    # d/dx* of 0.5||Gx||^2 = G^H G x = G(G x) for Hermitian G (torch grad = 2 dL/dx*... times 1)
    ref = pf.sparse_gram(y.detach(), G, device=cuda_device)
    assert rel(x.grad.cpu().numpy(), ref.cpu().numpy()) < 1e-5
