"""GPU parity of the Cartesian operators (fftn / ifftn / gram / fft / ifft)
against the CPU restatement.

Bar: relative L2 error <= 1e-5 (complex64 engine vs float64 restatement on
identical inputs); adjoint pairs to 1e-5 relative; the centred FFT utilities
are also checked against torch.fft with explicit shifts.
"""

import math
import zlib

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import restate as R
from pyfourier_amd import demo

pytestmark = pytest.mark.gpu

PARITY = 1e-5


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _mask(rng, nfr, shape, p=0.4):
    return (rng.random((nfr,) + tuple(shape)) < p).astype(np.float32)


def _basis(rng, T, K):
    return (demo.t2_subspace_basis(rng, T, K) * np.exp(1j * rng.random((T, K)))).astype(np.complex64)


CASES = [
    # name, shape, C, T (contrasts), mask frames (0 none), K (0 none), norm
    ("2d_plain", (64, 48), 3, 1, 0, 0, None),
    ("2d_frames_mask", (40, 36), 4, 5, 5, 0, "ortho"),
    ("2d_shared_mask", (32, 30), 6, 1, 1, 0, None),
    ("2d_basis", (24, 20), 2, 12, 12, 3, None),
    ("2d_basis_shared_mask", (30, 32), 3, 7, 1, 4, "ortho"),
    ("3d_mask", (16, 20, 24), 5, 1, 1, 0, "ortho"),
    ("3d_basis", (12, 10, 16), 2, 6, 6, 2, None),
    ("1d", (50,), 4, 1, 1, 0, None),
    ("odd_sizes", (45, 75), 2, 1, 1, 0, None),
    ("rocfft_sizes", (14, 22), 3, 2, 2, 0, "ortho"),
    ("rows_256", (256, 256), 2, 1, 1, 0, None),
    ("many_coils", (16, 64), 96, 1, 1, 0, None),
    ("coil_blocks", (8, 40), 130, 2, 2, 0, "ortho"),
    # 16 x Q-point register passes (column and row)
    ("q_sizes_2d", (160, 320), 3, 1, 1, 0, "ortho"),
    ("q_sizes_3d", (144, 24, 240), 2, 1, 1, 0, None),
    ("q_512", (512, 512), 2, 1, 1, 0, None),
    ("q_mask_frames", (192, 128), 2, 3, 3, 0, None),
    ("q_basis", (128, 160), 2, 6, 6, 2, None),
    ("q_640_400", (640, 400), 1, 1, 1, 0, "ortho"),
    ("q_288_576", (288, 576), 1, 1, 0, 0, None),
    ("q_80_96", (80, 96), 3, 1, 1, 0, "ortho"),
    ("q8_120_200", (120, 200), 2, 1, 1, 0, "ortho"),
    ("q8_360_3d", (360, 12, 120), 2, 2, 2, 0, None),
    ("q7_112_224", (112, 224), 2, 1, 1, 0, "ortho"),
    ("q7_336_280_448", (336, 280, 16), 1, 1, 1, 0, None),
    ("q7_448", (448, 20), 2, 1, 0, 0, None),
    ("q8_small", (40, 56, 72), 2, 1, 1, 0, "ortho"),
    ("q8_168", (168, 168), 2, 1, 1, 0, None),
]


@pytest.mark.parametrize("name,shape,C,T,mf,K,norm", CASES, ids=[c[0] for c in CASES])
def test_fftn_ifftn_parity(cuda_device, name, shape, C, T, mf, K, norm):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    d = len(shape)
    mask = _mask(rng, mf, shape) if mf else None
    basis = _basis(rng, T, K) if K else None
    nimg = K if K else T
    lead = (C, nimg) if (K or mf > 1) else (C,)
    x = demo.complex_gaussian(rng, lead + shape)
    plan = pf.plan_fft(mask, shape)
    y = pf.fftn(d, torch.from_numpy(x).to(cuda_device), plan=plan, basis=basis, norm=norm, device=cuda_device)
    xr = x.reshape((C, nimg) + shape)
    yr = R.cart_fft(xr, d, mask=mask, basis=basis, norm=norm)
    assert rel(y.cpu().numpy().reshape(yr.shape), yr) < PARITY, name
    w = demo.complex_gaussian(rng, yr.shape)
    wl = w.reshape(tuple(y.shape))
    xa = pf.ifftn(d, torch.from_numpy(wl).to(cuda_device), plan=plan, basis=basis, norm=norm, device=cuda_device)
    xar = R.cart_ifft(w, d, mask=mask, basis=basis, norm=norm)
    assert tuple(xa.shape) == x.shape
    assert rel(xa.cpu().numpy().reshape(xar.shape), xar) < PARITY, name


@pytest.mark.parametrize("K,L", [(0, 4), (3, 4), (8, 6)])
def test_fftn_b0_parity(cuda_device, K, L):
    """Field-map segments along the readout; (8, 6) exceeds one native call's
    32 segment-coefficient products and runs in segment chunks."""
    rng = np.random.default_rng(7 + K)
    shape = (32, 40)
    T = max(6, K + 2) if K else 1
    C = 3
    zmap = demo.smooth_field_map(rng, shape, 80.0)
    mask = _mask(rng, T, shape, 0.5)
    basis = _basis(rng, T, K) if K else None
    plan = pf.plan_fft(mask, shape, zmap=zmap, L=L, dt=4e-6)
    bt = plan.zmap_t_kernel.cpu().numpy()
    ct = plan.zmap_s_kernel.cpu().numpy()
    nimg = K if K else T
    x = demo.complex_gaussian(rng, (C, nimg) + shape)
    xt = torch.from_numpy(x if (K or T > 1) else x[:, 0]).to(cuda_device)
    y = pf.fftn(2, xt, plan=plan, basis=basis, device=cuda_device)
    yr = R.cart_fft(x, 2, mask=mask, basis=basis, b0=(bt, ct))
    assert rel(y.cpu().numpy().reshape(yr.shape), yr) < PARITY
    w = demo.complex_gaussian(rng, yr.shape)
    xa = pf.ifftn(2, torch.from_numpy(w.reshape(tuple(y.shape))).to(cuda_device), plan=plan, basis=basis,
                  device=cuda_device)
    xar = R.cart_ifft(w, 2, mask=mask, basis=basis, b0=(bt, ct))
    assert rel(xa.cpu().numpy().reshape(xar.shape), xar) < PARITY


def test_fftn_plans_from_mask_and_numpy_roundtrip(cuda_device):
    """No plan: fftn plans from `mask`; numpy in, numpy out; real input gives
    a real (real-part) output like the reference's dtype restore."""
    rng = np.random.default_rng(11)
    shape = (24, 32)
    mask = _mask(rng, 1, shape)[0].astype(np.int16)
    x = demo.complex_gaussian(rng, (2,) + shape)
    y = pf.fftn(2, x, mask=mask, device=cuda_device)
    assert isinstance(y, np.ndarray) and y.dtype == np.complex64
    assert rel(y, R.cart_fft(x[:, None], 2, mask=mask[None].astype(np.float32))[:, 0]) < PARITY
    xr = np.ascontiguousarray(x.real)
    yr = pf.fftn(2, xr, mask=mask, device=cuda_device)
    assert yr.dtype == np.float32
    assert rel(yr, R.cart_fft(xr[:, None], 2, mask=mask[None].astype(np.float32))[:, 0].real) < PARITY


@pytest.mark.parametrize("norm", [None, "ortho"])
def test_fftn_autograd_is_exact_adjoint(cuda_device, norm):
    rng = np.random.default_rng(12)
    shape = (20, 28)
    T, K, C = 5, 2, 3
    mask = _mask(rng, T, shape)
    basis = _basis(rng, T, K)
    plan = pf.plan_fft(mask, shape)
    x = demo.complex_gaussian(rng, (C, K) + shape)
    v = demo.complex_gaussian(rng, (C, T) + shape)
    xt = torch.from_numpy(x).to(cuda_device).requires_grad_()
    y = pf.fftn(2, xt, plan=plan, basis=basis, norm=norm, device=cuda_device)
    y.backward(gradient=torch.from_numpy(v).to(cuda_device))
    # exact adjoint of the forward in use: the ortho pair scaled back
    N = float(np.prod(shape))
    s = math.sqrt(N) if norm is None else 1.0
    want = R.cart_ifft(v, 2, mask=mask, basis=basis, norm="ortho") * s
    assert rel(xt.grad.cpu().numpy(), want) < PARITY
    vt = torch.from_numpy(v).to(cuda_device).requires_grad_()
    xa = pf.ifftn(2, vt, plan=plan, basis=basis, norm=norm, device=cuda_device)
    xa.backward(gradient=torch.from_numpy(x).to(cuda_device))
    s = 1.0 / math.sqrt(N) if norm is None else 1.0
    want = R.cart_fft(x, 2, mask=mask, basis=basis, norm="ortho") * s
    assert rel(vt.grad.cpu().numpy(), want) < PARITY


GRAM_CASES = [
    ("plain", (32, 24), 3, 1, 0, 0),
    ("frames", (20, 28), 2, 4, 4, 0),
    ("basis", (24, 24), 2, 8, 8, 3),
    ("b0", (16, 20), 3, 1, 1, 0),
    ("basis_b0", (16, 20), 2, 5, 5, 2),
    ("3d", (8, 12, 10), 2, 1, 1, 0),
]


@pytest.mark.parametrize("name,shape,C,T,mf,K", GRAM_CASES, ids=[c[0] for c in GRAM_CASES])
def test_gram_parity(cuda_device, name, shape, C, T, mf, K):
    rng = np.random.default_rng(20 + len(name))
    d = len(shape)
    mask = _mask(rng, mf, shape, 0.5) if mf else None
    basis = _basis(rng, T, K) if K else None
    zmap = tm = None
    if "b0" in name:
        zmap = demo.smooth_field_map(rng, shape, 60.0)
        tm = np.linspace(0, 2e-3, shape[-1])
    gm = pf.plan_gram(mask, shape, basis=basis, zmap=zmap, L=3, T=tm, device=cuda_device)
    nimg = K if K else T
    lead = (C, nimg) if (K or mf > 1) else (C,)
    x = demo.complex_gaussian(rng, lead + shape)
    g = pf.gram(torch.from_numpy(x).to(cuda_device), gm, device=cuda_device)
    b0 = None
    if zmap is not None:
        b0 = (gm.zmap_t_kernel.cpu().numpy(), gm.zmap_s_kernel.cpu().numpy())
    xr = x.reshape((C, nimg) + shape)
    gr = R.cart_gram(xr, shape, nframes=nimg if basis is None else 1, mask=mask, basis=basis, b0=b0)
    assert rel(g.cpu().numpy().reshape(gr.shape), gr) < PARITY, name
    # the normal operator of the masked FFT (binary mask)
    plan = pf.plan_fft(mask, shape, zmap=zmap, L=3, T=tm)
    ref = pf.ifftn(d, pf.fftn(d, torch.from_numpy(x).to(cuda_device), plan=plan, basis=basis, device=cuda_device),
                   plan=plan, basis=basis, device=cuda_device)
    assert rel(g.cpu().numpy(), ref.cpu().numpy()) < 2 * PARITY, name


def test_gram_autograd_self_adjoint(cuda_device):
    rng = np.random.default_rng(30)
    shape = (24, 20)
    mask = _mask(rng, 1, shape)
    gm = pf.plan_gram(mask, shape, device=cuda_device)
    x = torch.from_numpy(demo.complex_gaussian(rng, (2,) + shape)).to(cuda_device).requires_grad_()
    v = torch.from_numpy(demo.complex_gaussian(rng, (2,) + shape)).to(cuda_device)
    pf.gram(x, gm, device=cuda_device).backward(gradient=v)
    want = pf.gram(v, gm, device=cuda_device)
    assert rel(x.grad.cpu().numpy(), want.cpu().numpy()) < PARITY


@pytest.mark.parametrize("axes,norm,centered", [(None, "ortho", True), ((-2, -1), None, True), ((0, 2), "forward", True),
                                                ((-1,), "ortho", False), ((1, 2, 3), "backward", False)])
def test_fft_utilities_vs_torch(cuda_device, axes, norm, centered):
    rng = np.random.default_rng(31)
    x = torch.from_numpy(demo.complex_gaussian(rng, (3, 12, 10, 8))).to(cuda_device)
    if axes is None:
        x = x[0]
    ax = tuple(range(x.dim())) if axes is None else axes
    tn = "backward" if norm is None else norm
    if centered:
        ref = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(x, dim=ax), dim=ax, norm=tn), dim=ax)
        iref = torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(x, dim=ax), dim=ax, norm=tn), dim=ax)
    else:
        ref = torch.fft.fftn(x, dim=ax, norm=tn)
        iref = torch.fft.ifftn(x, dim=ax, norm=tn)
    y = pf.fft(x, axes=axes, norm=norm, centered=centered)
    yi = pf.ifft(x, axes=axes, norm=norm, centered=centered)
    assert rel(y.cpu().numpy(), ref.cpu().numpy()) < PARITY
    assert rel(yi.cpu().numpy(), iref.cpu().numpy()) < PARITY


def test_fft_utility_numpy_and_inverse(cuda_device):
    rng = np.random.default_rng(32)
    x = demo.complex_gaussian(rng, (2, 30, 18))
    k = pf.fft(x, axes=(-2, -1))
    assert isinstance(k, np.ndarray)
    assert rel(k, R.fftc(x.astype(np.complex128), 2, norm="ortho")) < PARITY
    assert rel(pf.ifft(k, axes=(-2, -1)), x) < PARITY
