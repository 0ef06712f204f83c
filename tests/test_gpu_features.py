"""GPU parity of the operator variants against the CPU restatement:
1-D, anisotropic 3-D, Cartesian axes, rocFFT fallback grid sizes,
multi-frame trajectories, the fused low-rank subspace, the time-segmented
B0 model, density weights, real inputs and numpy inputs.
"""

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
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check_pair(dev, coord, shape, x, width=4, oversamp=2.0, **kw):
    P = R.Plan(coord, shape, width=width, oversamp=oversamp)
    plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=width, oversamp=oversamp)
    y = pf.nufft(torch.from_numpy(x).to(dev), nufft_plan=plan, device=dev, **kw)
    F = P.nframes
    B = x.shape[0] if x.ndim > len(P.shape) + (1 if F > 1 else 0) else 1
    xr = x.reshape((B, F) + tuple(P.shape))
    yr = R.nufft(xr, P, norm=kw.get("norm"))
    assert rel(y.cpu().numpy(), yr) < PARITY
    ya = demo.complex_gaussian(np.random.default_rng(9), tuple(y.shape))
    xa = pf.nufft_adj(torch.from_numpy(ya).to(dev), nufft_plan=plan, device=dev, **kw)
    xar = R.nufft_adj(ya.reshape(B, F, -1), P, norm=kw.get("norm"))
    assert rel(xa.cpu().numpy(), xar) < PARITY
    return plan


def test_1d(cuda_device):
    rng = np.random.default_rng(0)
    coord = np.sort(rng.uniform(-np.pi, np.pi, size=(3000, 1))).astype(np.float32)
    coord[0, 0] = -np.pi
    x = demo.complex_gaussian(rng, (4, 200))
    _check_pair(cuda_device, coord, 200, x)


def test_3d_anisotropic(cuda_device):
    rng = np.random.default_rng(1)
    coord = demo.golden_means_radial_3d(200, 48)
    shape = (20, 24, 32)
    x = demo.complex_gaussian(rng, (3,) + shape)
    _check_pair(cuda_device, coord, shape, x, width=(4, 5, 6), oversamp=(2.0, 1.5, 1.25))


def test_rocfft_fallback_size(cuda_device):
    # oversampled size 147 = 3 * 7^2 is outside the fused FFT's radix set
    rng = np.random.default_rng(2)
    coord = demo.golden_angle_radial_2d(40, 120)
    x = demo.complex_gaussian(rng, (2, 98, 98))
    _check_pair(cuda_device, coord, 98, x, oversamp=1.5, norm="ortho")


def test_cartesian_axis_stack_of_stars(cuda_device):
    # z on integer grid positions -> width 1, oversamp 1 on that axis
    rng = np.random.default_rng(3)
    Nz, N = 8, 32
    c2 = demo.golden_angle_radial_2d(24, 64) / np.pi * (N / 2)
    kz = np.arange(Nz) - Nz // 2
    coord = np.zeros((Nz, 24, 64, 3), dtype=np.float32)
    coord[..., :2] = c2[None]
    coord[..., 2] = kz[:, None, None]
    coord = coord.reshape(Nz * 24, 64, 3)
    plan = pf.plan_nufft(torch.from_numpy(coord), (Nz, N, N), width=4, oversamp=2.0)
    assert plan.width[0] == 1 and plan.oversamp[0] == 1.0
    x = demo.complex_gaussian(rng, (2, Nz, N, N))
    _check_pair(cuda_device, coord, (Nz, N, N), x)


def test_cartesian_exact_fft(cuda_device):
    # integer coordinates with W=1, sigma=1 reduce exactly to a centred FFT
    rng = np.random.default_rng(4)
    N = 32
    ky, kx = np.meshgrid(np.arange(N) - N // 2, np.arange(N) - N // 2, indexing="ij")
    coord = np.stack([kx, ky], -1).reshape(N, N, 2).astype(np.float32)
    x = demo.complex_gaussian(rng, (N, N))
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), torch.from_numpy(coord), shape=N, device=cuda_device)
    ref = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x)))
    assert rel(y.cpu().numpy(), ref) < 1e-6


def test_multiframe(cuda_device):
    rng = np.random.default_rng(5)
    F, N = 3, 48
    coord = np.stack([demo.golden_angle_radial_2d(16, 96) * (1.0 - 0.1 * f) for f in range(F)])
    coord[:, 0, 0] = [-np.pi, 0]
    x = demo.complex_gaussian(rng, (2, F, N, N))
    _check_pair(cuda_device, coord, N, x)


def test_subspace(cuda_device):
    rng = np.random.default_rng(6)
    F, K, N, C = 20, 4, 48, 3
    coord = np.stack([demo.golden_angle_radial_2d(1, 128)[0] for _ in range(F)])[:, None]
    ang = np.arange(F) * 2.4
    rot = np.stack([np.cos(ang), -np.sin(ang), np.sin(ang), np.cos(ang)], -1).reshape(F, 1, 1, 2, 2)
    coord = np.einsum("fvsij,fvsj->fvsi", np.broadcast_to(rot, coord.shape[:3] + (2, 2)), coord).astype(np.float32)
    basis = demo.t2_subspace_basis(rng, F, K)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0)
    x = demo.complex_gaussian(rng, (C, K, N, N))
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, basis=torch.from_numpy(basis),
                 device=cuda_device)
    assert tuple(y.shape) == (C, F, 1, 128)
    yr = R.nufft(x, P, basis=basis)
    assert rel(y.cpu().numpy().reshape(C, F, -1), yr) < PARITY
    ya = demo.complex_gaussian(rng, (C, F, 1, 128))
    xa = pf.nufft_adj(torch.from_numpy(ya).to(cuda_device), nufft_plan=plan, basis=torch.from_numpy(basis),
                      device=cuda_device)
    assert tuple(xa.shape) == (C, K, N, N)
    xar = R.nufft_adj(ya.reshape(C, F, -1), P, basis=basis)
    assert rel(xa.cpu().numpy(), xar) < PARITY


def test_b0_segments(cuda_device):
    rng = np.random.default_rng(7)
    N, C, L = 48, 2, 6
    coord = demo.variable_density_spiral_2d(8, 512, N)
    zmap = demo.smooth_field_map(rng, (N, N), 100.0)
    t = np.linspace(0, 10e-3, 512).astype(np.float32)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0, zmap=torch.from_numpy(zmap),
                         T=torch.from_numpy(t), L=L, nbins=40)
    bt = plan.zmap_t_kernel.cpu().numpy().astype(np.complex128)
    ct = plan.zmap_s_kernel.cpu().numpy().astype(np.complex128)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    x = demo.complex_gaussian(rng, (C, N, N))
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    yr = R.nufft(x[:, None], P, b0=(bt, ct))
    assert rel(y.cpu().numpy().reshape(C, 1, -1), yr) < PARITY
    # the segmented model approximates the exact field-map signal
    ye = U.nudft_b0(x[:1], P.coord.reshape(-1, 2), (N, N), zmap, t)
    assert rel(y.cpu().numpy()[:1].reshape(1, -1), ye) < 2e-2
    ya = demo.complex_gaussian(rng, tuple(y.shape))
    xa = pf.nufft_adj(torch.from_numpy(ya).to(cuda_device), nufft_plan=plan, device=cuda_device)
    xar = R.nufft_adj(ya.reshape(C, 1, -1), P, b0=(bt, ct))
    assert rel(xa.cpu().numpy().reshape(C, 1, N, N), xar) < PARITY
    # segment chunking (L_batch_size) gives the same operator
    plan.zmap_batch_size = 4
    y2 = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y2.cpu().numpy(), y.cpu().numpy()) < 1e-6


def test_b0_segments_on_a_7_smooth_grid(cuda_device):
    """224-point rows (2^5 x 7) have register FFT passes only: up to 16
    segments per call run fused, so the plan batches 20 segments by 16 (the
    operator matches the oracle); an explicit batch of 20 raises a clear
    error."""
    rng = np.random.default_rng(9)
    N, C, L = 112, 2, 20
    coord = demo.variable_density_spiral_2d(4, 512, N)
    zmap = demo.smooth_field_map(rng, (N, N), 100.0)
    t = np.linspace(0, 8e-3, 512).astype(np.float32)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=4, oversamp=2.0, zmap=torch.from_numpy(zmap),
                         T=torch.from_numpy(t), L=L, nbins=40)
    assert plan.interpolator.os_shape == (224, 224)
    b0 = (plan.zmap_t_kernel.cpu().numpy().astype(np.complex128), plan.zmap_s_kernel.cpu().numpy().astype(np.complex128))
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    x = demo.complex_gaussian(rng, (C, N, N))
    assert plan.zmap_batch_size == 16
    plan.zmap_batch_size = 20
    with pytest.raises(RuntimeError, match="segment batch size"):
        pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    plan.zmap_batch_size = 16
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y.cpu().numpy().reshape(C, 1, -1), R.nufft(x[:, None], P, b0=b0)) < PARITY
    ya = demo.complex_gaussian(rng, tuple(y.shape))
    xa = pf.nufft_adj(torch.from_numpy(ya).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(xa.cpu().numpy().reshape(C, 1, N, N), R.nufft_adj(ya.reshape(C, 1, -1), P, b0=b0)) < PARITY


def test_weights_real_and_numpy(cuda_device):
    rng = np.random.default_rng(8)
    N = 40
    coord = demo.golden_angle_radial_2d(20, 80)
    w = demo.spiral_density(coord)
    P = R.Plan(coord, N, width=4, oversamp=1.25)  # nufft's default oversampling
    xr = rng.standard_normal((N, N)).astype(np.float32)
    y = pf.nufft(xr, coord, shape=N, weight=w, device="cuda:0")
    assert isinstance(y, np.ndarray) and y.dtype == np.float32
    yr = R.nufft(xr[None, None].astype(np.complex128), P, weight=w)
    assert rel(y, yr.real.reshape(y.shape)) < PARITY
    yc = demo.complex_gaussian(rng, coord.shape[:-1])
    xa = pf.nufft_adj(yc, coord, shape=N, weight=w, device="cuda:0")
    assert isinstance(xa, np.ndarray) and xa.dtype == np.complex64
    xar = R.nufft_adj(yc.reshape(1, 1, -1), P, weight=w)
    assert rel(xa, xar) < PARITY


def test_cpu_device_raises():
    coord = demo.golden_angle_radial_2d(4, 16)
    with pytest.raises(RuntimeError, match="GPU only"):
        pf.nufft(np.zeros((8, 8), np.complex64), coord, shape=8, device="cpu")


@pytest.mark.parametrize("shape,oversamp", [((6, 6, 6), 2.0), ((17, 17), 2.0), ((18, 18), 2.0), ((5, 40), 2.0)])
def test_small_grids_periodic_wrap(cuda_device, shape, oversamp):
    # grids of one spread tile (footprints wrap inside the tile) and grids
    # just above a tile, where footprints cross the periodic boundary
    rng = np.random.default_rng(10)
    d = len(shape)
    coord = rng.uniform(-np.pi, np.pi, size=(400, d)).astype(np.float32)
    coord[0] = 0
    coord[0, 0] = -np.pi
    x = demo.complex_gaussian(rng, (3,) + shape)
    _check_pair(cuda_device, coord, shape, x, oversamp=oversamp)


@pytest.mark.parametrize("shape,oversamp", [((12, 128), 2.0), ((12, 128), 1.25), ((12, 192), 2.0), ((12, 384), 1.25),
                                           ((12, 160), 1.25)])
def test_subspace_and_frames_register_rows(cuda_device, shape, oversamp):
    """x axes of 256, 160, 384 and 480 points run the register row FFTs
    (16 x 16, 16 x 10, 16 x 24, 16 x 30), with a subspace basis (image planes
    = coefficients) and with per-frame images."""
    rng = np.random.default_rng(21)
    F, K, C = 3, 2, 3
    coord = (rng.uniform(-0.5, 0.5, (F, 2, 150, 2)) * np.array(shape[::-1])).astype(np.float32)
    coord[0, 0, 0] = [-shape[1] // 2, -6]
    P = R.Plan(coord, shape, width=4, oversamp=oversamp)
    plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=4, oversamp=oversamp)
    basis = demo.t2_subspace_basis(rng, F, K)
    x = demo.complex_gaussian(rng, (C, K) + shape)
    y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, basis=torch.from_numpy(basis),
                 device=cuda_device)
    assert rel(y.cpu().numpy().reshape(C, F, -1), R.nufft(x, P, basis=basis)) < PARITY
    ya = demo.complex_gaussian(rng, (C, F) + coord.shape[1:-1])
    xa = pf.nufft_adj(torch.from_numpy(ya).to(cuda_device), nufft_plan=plan, basis=torch.from_numpy(basis),
                      device=cuda_device)
    assert rel(xa.cpu().numpy(), R.nufft_adj(ya.reshape(C, F, -1), P, basis=basis)) < PARITY
    xf = demo.complex_gaussian(rng, (C, F) + shape)
    yf = pf.nufft(torch.from_numpy(xf).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(yf.cpu().numpy().reshape(C, F, -1), R.nufft(xf, P)) < PARITY
    xfa = pf.nufft_adj(torch.from_numpy(ya).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(xfa.cpu().numpy(), R.nufft_adj(ya.reshape(C, F, -1), P)) < PARITY


@pytest.mark.parametrize("shape,oversamp", [((10, 256), 2.0), ((10, 128), 1.25), ((10, 256), 1.25), ((10, 192), 2.0),
                                           ((10, 160), 1.25)])
def test_b0_with_frames_and_subspace_register_rows(cuda_device, shape, oversamp):
    """x axes of 512, 160, 320 and 384 points run the register row FFTs with
    B0 segments, per frame and with a subspace basis.  3 coils: the forward's
    16-column blocks span grid rows (15 or 30 values per grid point); with
    the basis the adjoint splits the 6 (k, coil) pairs over two blocks, each
    summing all 5 segments."""
    rng = np.random.default_rng(22)
    F, K, C, L = 3, 2, 3, 5
    coord = (rng.uniform(-0.5, 0.5, (F, 2, 150, 2)) * np.array(shape[::-1])).astype(np.float32)
    coord[0, 0, 0] = [-shape[1] // 2, -5]
    zmap = demo.smooth_field_map(rng, shape, 100.0)
    t = np.linspace(0, 8e-3, 150).astype(np.float32)
    plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=4, oversamp=oversamp, zmap=torch.from_numpy(zmap),
                         T=torch.from_numpy(t), L=L, nbins=40)
    b0 = (plan.zmap_t_kernel.cpu().numpy().astype(np.complex128),
          plan.zmap_s_kernel.cpu().numpy().astype(np.complex128))
    P = R.Plan(coord, shape, width=4, oversamp=oversamp)
    basis = demo.t2_subspace_basis(rng, F, K)
    for bas in (None, basis):
        tb = None if bas is None else torch.from_numpy(bas)
        x = demo.complex_gaussian(rng, (C, F if bas is None else K) + shape)
        y = pf.nufft(torch.from_numpy(x).to(cuda_device), nufft_plan=plan, basis=tb, device=cuda_device)
        assert rel(y.cpu().numpy().reshape(C, F, -1), R.nufft(x, P, basis=bas, b0=b0)) < PARITY
        ya = demo.complex_gaussian(rng, (C, F) + coord.shape[1:-1])
        xa = pf.nufft_adj(torch.from_numpy(ya).to(cuda_device), nufft_plan=plan, basis=tb, device=cuda_device)
        assert rel(xa.cpu().numpy(), R.nufft_adj(ya.reshape(C, F, -1), P, basis=bas, b0=b0)) < PARITY


def _expanded_case(cuda_device, shape, coord, C, L=1, basis=None, nt=None, seed=0):
    """Adjoint of a B0 / subspace batch whose R*C values form whole 32-value
    chunks: the sort writes them expanded and the spread runs the plain
    pipelined (2-D) or matrix-core (3-D) kernel.  Checked against the oracle
    and against the staged path (sort + on-the-fly spread, k_spread_reg)."""
    from pyfourier_amd._nufft import _Call, stage_timer

    rng = np.random.default_rng(seed)
    kw = {}
    if L > 1:
        zmap = demo.smooth_field_map(rng, shape, 80.0)
        kw = dict(zmap=torch.from_numpy(zmap), T=torch.from_numpy(np.linspace(0, 6e-3, nt).astype(np.float32)),
                  L=L, nbins=40)
    plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=4, oversamp=2.0, **kw)
    P = R.Plan(coord, shape, width=4, oversamp=2.0)
    b0 = None
    if L > 1:
        b0 = (plan.zmap_t_kernel.cpu().numpy().astype(np.complex128),
              plan.zmap_s_kernel.cpu().numpy().astype(np.complex128))
    tb = None if basis is None else torch.from_numpy(basis)
    call = _Call(plan, cuda_device, tb, None, None)
    what = call.describe(C, call.segments[0])
    F = coord.shape[0] if basis is not None else 1
    ya = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    yt = torch.from_numpy(ya).to(cuda_device)
    xa = pf.nufft_adj(yt, nufft_plan=plan, basis=tb, device=cuda_device)
    xr = R.nufft_adj(ya.reshape(C, F, -1), P, basis=basis, b0=b0)
    e = rel(xa.cpu().numpy().reshape(xr.shape), xr)
    with stage_timer():
        xs = pf.nufft_adj(yt, nufft_plan=plan, basis=tb, device=cuda_device)
    torch.cuda.synchronize(cuda_device)
    es = rel(xa.cpu().numpy(), xs.cpu().numpy())
    print(f"  {what}: oracle {e:.2e}, staged {es:.2e}", flush=True)
    assert e < PARITY and es < 1e-6
    return what


def test_expanded_adjoint_b0_2d(cuda_device):
    shape = (64, 64)
    coord = demo.variable_density_spiral_2d(8, 512, 64)
    what = _expanded_case(cuda_device, shape, coord, C=8, L=4, nt=512, seed=31)
    assert "spread=k_sort_x+k_spread_pipe<" in what, what


def test_expanded_adjoint_subspace_2d(cuda_device):
    rng = np.random.default_rng(32)
    F, K, N = 6, 4, 48
    coord = np.stack([demo.golden_angle_radial_2d(2, 96) for _ in range(F)])
    coord = (coord * rng.uniform(0.8, 1.0, (F, 1, 1, 1))).astype(np.float32)
    basis = demo.t2_subspace_basis(rng, F, K)
    what = _expanded_case(cuda_device, (N, N), coord, C=8, basis=basis, seed=33)
    assert "spread=k_sort_x+k_spread_pipe<" in what, what


def test_expanded_adjoint_b0_3d_matrix_core(cuda_device):
    shape = (24, 24, 24)
    coord = demo.golden_means_radial_3d(256, 48)
    what = _expanded_case(cuda_device, shape, coord, C=8, L=4, nt=48, seed=34)
    assert "spread=k_sort_x+k_tile_max+k_spread_mfma<" in what, what
