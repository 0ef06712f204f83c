"""GPU tests of the Toeplitz Gram (plan_spgram / sparse_gram).

The Gram must equal the oracle's Toeplitz apply with FP64-exact PSFs (the
plain kernel, or the per-voxel R x R block of the subspace and field-map
Grams) up to the PSF's own NUFFT error, and the normal operator
nufft_adj(w * nufft(x)) of the NUFFT it replaces up to kernel-approximation
error.
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
    xh = demo.complex_gaussian(rng, (C, K, N, N))
    x = torch.from_numpy(xh).to(cuda_device)
    gx = pf.sparse_gram(x, G, device=cuda_device)
    # oracle: R x R block Toeplitz (per-voxel K x K, _bmatvec/_bdot.py:130-157)
    # with the FP64-exact PSFs, c_k(j) = basis[frame(j), k]
    P = R.Plan(coord, N, width=6, oversamp=2.0)
    M = coord.shape[1] * coord.shape[2]
    coefs = basis[np.repeat(np.arange(F), M)]
    T = R.toeplitz_kernel_block_exact(P.coord.reshape(-1, 2), (N, N), coefs, ghat=float(np.prod(P.os_shape)))
    ref = np.stack([R.toeplitz_apply_block(xh[c], T, (N, N)) for c in range(C)])
    assert rel(gx.cpu().numpy(), ref) < 1e-4
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
    xh = demo.complex_gaussian(rng, (C, N, N))
    x = torch.from_numpy(xh).to(cuda_device)
    gx = pf.sparse_gram(x, G, device=cuda_device)
    # oracle: R x R block Toeplitz over the L segments (_bdot.py:143-157 with
    # the segment loop of _sparsegram.py:26-53), FP64-exact PSFs, c_l(j) = Bt[l, t(j)]
    bt, ct = pf.mri_exp_approx(torch.from_numpy(zmap), torch.from_numpy(t), L, (40, 40))
    bt, ct = bt.numpy(), ct.numpy()
    P = R.Plan(coord, N, width=6, oversamp=2.0)
    M = coord.shape[-3] * coord.shape[-2]
    coefs = bt[:, np.arange(M) % coord.shape[-2]].T
    T = R.toeplitz_kernel_block_exact(P.coord.reshape(-1, 2), (N, N), coefs, ghat=float(np.prod(P.os_shape)))
    ref = np.stack([R.toeplitz_apply_block(xh[c][None], T, (N, N), ct)[0] for c in range(C)])
    assert rel(gx.cpu().numpy(), ref) < 1e-4
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0, zmap=torch.from_numpy(zmap),
                         T=torch.from_numpy(t), L=L, nbins=40)
    ex = pf.nufft_adj(pf.nufft(x, nufft_plan=plan, device=cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(gx.cpu().numpy(), ex.cpu().numpy()) < 2e-4


def test_gram_subspace_b0_generic_block(cuda_device):
    """Subspace basis and field map together: R = L * K = 24 > 16, the
    block size beyond the unrolled instances (LDS-staged product)."""
    rng = np.random.default_rng(5)
    F, K, L, N, C, ns = 6, 4, 6, 16, 2, 96
    base = demo.variable_density_spiral_2d(1, ns, N)[0]
    coord = np.stack([base @ np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]]).T
                      for a in 2.39996 * np.arange(F)])[:, None].astype(np.float32)  # (F, 1, ns, 2)
    basis = demo.t2_subspace_basis(rng, F, K).astype(np.complex64)
    zmap = demo.smooth_field_map(rng, (N, N), 60.0)
    t = np.linspace(0, 5e-3, ns).astype(np.float32)
    G = pf.plan_spgram(torch.from_numpy(coord), N, basis=torch.from_numpy(basis), zmap=torch.from_numpy(zmap),
                       T=torch.from_numpy(t), L=L, nbins=40, oversamp=2.0, device=cuda_device)
    xh = demo.complex_gaussian(rng, (C, K, N, N))
    gx = pf.sparse_gram(torch.from_numpy(xh).to(cuda_device), G, device=cuda_device).cpu().numpy()
    bt, ct = pf.mri_exp_approx(torch.from_numpy(zmap), torch.from_numpy(t), L, (40, 40))
    bt, ct = bt.numpy(), ct.numpy()
    P = R.Plan(coord, N, width=6, oversamp=2.0)
    M = F * ns
    coefs = (bt[:, np.arange(M) % ns].T[:, :, None] * basis[np.arange(M) // ns][:, None, :]).reshape(M, L * K)
    T = R.toeplitz_kernel_block_exact(P.coord.reshape(-1, 2), (N, N), coefs, ghat=float(np.prod(P.os_shape)))
    ref = np.stack([R.toeplitz_apply_block(xh[c], T, (N, N), ct) for c in range(C)])
    assert rel(gx, ref) < 1e-4


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
