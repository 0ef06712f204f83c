"""GPU parity of the sparse FFT (plan_spfft / sparse_fftn / sparse_ifftn) and
of GRAPPA operator gridding (grog_interp) against the CPU restatement.

Bar: relative L2 error <= 1e-5 for complex data (complex64 engine vs
float64 restatement on identical inputs); indexes and repeat-count weights
exactly equal.
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
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _indexes(rng, F, nv, ns, shape_zyx, dup=True):
    d = len(shape_zyx)
    cols = [rng.integers(0, shape_zyx[d - 1 - a], (F, nv, ns)) for a in range(d)]  # x, y, z
    idx = np.stack(cols, -1)
    if dup:
        idx[:, :, 1] = idx[:, :, 0]  # repeated points
    return idx


SPARSE_CASES = [
    # name, shape, C, F, K, norm, weighted
    ("2d", (48, 40), 3, 1, 0, None, False),
    ("2d_frames_ortho_weighted", (32, 36), 2, 4, 0, "ortho", True),
    ("2d_basis", (30, 32), 2, 6, 3, None, True),
    ("3d", (16, 20, 18), 2, 1, 0, "ortho", False),
    ("3d_basis", (12, 16, 10), 2, 5, 2, None, False),
]


@pytest.mark.parametrize("name,shape,C,F,K,norm,weighted", SPARSE_CASES, ids=[c[0] for c in SPARSE_CASES])
def test_sparse_fft_parity(cuda_device, name, shape, C, F, K, norm, weighted):
    rng = np.random.default_rng(len(name) * 7 + C)
    nv, ns = 6, 50
    idx = _indexes(rng, F, nv, ns, shape)
    basis = (demo.t2_subspace_basis(rng, F, K) * np.exp(1j * rng.random((F, K)))).astype(np.complex64) if K else None
    w = rng.uniform(0.5, 2.0, (F, nv, ns)).astype(np.float32) if weighted else None
    nimg = K if K else F
    lead = (C, nimg) if nimg > 1 else (C,)
    x = demo.complex_gaussian(rng, lead + shape)
    plan = pf.plan_spfft(idx, shape)
    y = pf.sparse_fftn(torch.from_numpy(x).to(cuda_device), plan=plan, basis=basis, weight=w, norm=norm,
                       device=cuda_device)
    yr = R.sparse_fft(x.reshape((C, nimg) + shape), shape, idx.reshape(F, -1, len(shape)), basis=basis,
                      weight=w, norm=norm)
    assert rel(y.cpu().numpy().reshape(yr.shape), yr) < PARITY, name
    v = demo.complex_gaussian(rng, yr.shape)
    xa = pf.sparse_ifftn(torch.from_numpy(v.reshape(tuple(y.shape))).to(cuda_device), plan=plan, basis=basis,
                         weight=w, norm=norm, device=cuda_device)
    xr = R.sparse_ifft(v, shape, idx.reshape(F, -1, len(shape)), basis=basis, weight=w, norm=norm)
    assert rel(xa.cpu().numpy().reshape(xr.shape), xr) < PARITY, name


def test_sparse_fft_b0_and_centred_indexes(cuda_device):
    rng = np.random.default_rng(60)
    shape = (32, 40)
    nv, ns = 8, 64
    idx = _indexes(rng, 1, nv, ns, shape)
    zmap = demo.smooth_field_map(rng, shape, 70.0)
    plan = pf.plan_spfft(idx, shape, zmap=zmap, L=4, dt=5e-6)
    x = demo.complex_gaussian(rng, (2,) + shape)
    y = pf.sparse_fftn(torch.from_numpy(x).to(cuda_device), plan=plan, device=cuda_device)
    b0 = (plan.zmap_t_kernel.cpu().numpy(), plan.zmap_s_kernel.cpu().numpy())
    yr = R.sparse_fft(x[:, None], shape, idx.reshape(1, -1, 2), b0=b0)
    assert rel(y.cpu().numpy().reshape(yr.shape), yr) < PARITY
    v = demo.complex_gaussian(rng, yr.shape)
    xa = pf.sparse_ifftn(torch.from_numpy(v.reshape(tuple(y.shape))).to(cuda_device), plan=plan, device=cuda_device)
    xr = R.sparse_ifft(v, shape, idx.reshape(1, -1, 2), b0=b0)
    assert rel(xa.cpu().numpy().reshape(xr.shape), xr) < PARITY
    # centred indexes (negative values) name the same points as [0, N) indexes
    cen = idx - np.array([shape[1] // 2, shape[0] // 2])
    y2 = pf.sparse_fftn(torch.from_numpy(x).to(cuda_device), cen, device=cuda_device)
    y1 = pf.sparse_fftn(torch.from_numpy(x).to(cuda_device), idx, device=cuda_device)
    assert torch.equal(y1, y2)


def test_sparse_fft_matches_masked_fftn(cuda_device):
    """Picking the points of a mask equals the masked dense FFT at those points."""
    rng = np.random.default_rng(61)
    shape = (24, 32)
    mask = rng.random(shape) < 0.3
    yy, xx = np.nonzero(mask)
    idx = np.stack([xx, yy], -1)[None, None]  # (1, 1, M, 2)
    x = torch.from_numpy(demo.complex_gaussian(rng, (3,) + shape)).to(cuda_device)
    ys = pf.sparse_fftn(x, idx, device=cuda_device)
    yd = pf.fftn(2, x, mask=mask.astype(np.float32), device=cuda_device)
    assert rel(ys.cpu().numpy().reshape(3, -1), yd.cpu().numpy()[:, yy, xx]) < PARITY


GROG_CASES = [
    # name, ndim, nc, calib slices, data batch, nsteps
    ("2d_per_slice", 2, 8, 3, 3, 11),
    ("2d_shared", 2, 6, 1, 2, 11),
    ("2d_nc32_lds", 2, 32, 1, 1, 11),
    ("2d_nc40", 2, 40, 1, 1, 11),
    ("2d_nc32_21steps_global_tables", 2, 32, 1, 1, 21),
    ("3d", 3, 12, 1, 1, 11),
]


@pytest.mark.parametrize("name,ndim,nc,nsl,B,nsteps", GROG_CASES, ids=[c[0] for c in GROG_CASES])
def test_grog_parity(cuda_device, name, ndim, nc, nsl, B, nsteps):
    rng = np.random.default_rng(len(name) + nc)
    n = 24 if ndim == 2 else 16
    cal_n = 12 if ndim == 2 else 8
    if ndim == 2:
        calib = demo.complex_gaussian(rng, (nc, nsl, cal_n, cal_n)) if nsl > 1 else \
            demo.complex_gaussian(rng, (nc, cal_n, cal_n))
    else:
        calib = demo.complex_gaussian(rng, (nc, cal_n, cal_n, cal_n))
    F, nv, ns = 2, 6, 40
    coord = rng.uniform(-n / 2 - 0.6, n / 2 + 0.4, (F, nv, ns, ndim)).astype(np.float32)
    x = demo.complex_gaussian(rng, (B, nc, F, nv, ns))
    out, idx, w = pf.grog_interp(torch.from_numpy(x).to(cuda_device), calib, coord, n, lamda=0.01, nsteps=nsteps)
    ro, ri, rw = R.grog_interp(x.reshape(B, nc, -1), calib, coord.reshape(-1, ndim), (n,) * ndim, lamda=0.01,
                               nsteps=nsteps)
    assert tuple(out.shape) == x.shape and tuple(idx.shape) == coord.shape and tuple(w.shape) == coord.shape[:-1]
    np.testing.assert_array_equal(idx.cpu().numpy().reshape(-1, ndim), ri)
    np.testing.assert_array_equal(w.cpu().numpy().ravel(), rw)
    assert rel(out.cpu().numpy().reshape(ro.shape), ro) < PARITY, name


def test_grog_numpy_in_numpy_out(cuda_device):
    rng = np.random.default_rng(62)
    nc = 4
    calib = demo.complex_gaussian(rng, (nc, 10, 10))
    coord = rng.uniform(-7.5, 7.5, (1, 5, 16, 2)).astype(np.float32)
    x = demo.complex_gaussian(rng, (nc, 1, 5, 16))
    out, idx, w = pf.grog_interp(x, calib, coord, 16)
    assert isinstance(out, np.ndarray) and out.shape == x.shape and idx.dtype == np.int64
    ro, ri, rw = R.grog_interp(x.reshape(1, nc, -1), calib, coord.reshape(-1, 2), (16, 16))
    assert rel(out.reshape(ro.shape), ro) < PARITY
