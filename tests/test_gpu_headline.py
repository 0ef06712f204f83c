"""The headline workload's own kernel instances against the oracle.

The 3-D multicoil headline (golden-means radial, 32 coils batched, sigma 2,
|k|^2 density weights, KB width 6 and the configured width 4) runs the
LDS-staged gather k_gather_col<W> and the matrix-core 32-coil spread
k_spread_mfma<8, 4, 4>.  Here they are checked:

* on a reduced grid (32^3 image, 64^3 grid) against the float64 restatement
  oracle.restate (src/pyfourier/_subroutines/_grid/_cart2noncart/_3D.py:13-51,
  _noncart2cart/_3D.py:13-51 and the pipeline of _nufft/_nufft.py:218-287),
  forward and adjoint, relative L2 <= 1e-5;
* at the full 128^3 x 32-coil size through size-independent properties:
  ortho adjointness, bitwise run-to-run determinism of the adjoint, and the
  width-6 forward's error against the FP64 NUDFT on a sample subset < 1e-4.
"""

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import nudft as U
from oracle import restate as R
from pyfourier_amd import demo
from pyfourier_amd._nufft import _Call

pytestmark = pytest.mark.gpu

PARITY = 1e-5


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _ramp(coord):
    k = np.sqrt((coord.astype(np.float64) ** 2).sum(-1))
    return ((k / k.max()) ** 2).astype(np.float32)


@pytest.mark.parametrize("width,nspokes", [(4, 1024), (6, 512)])
def test_headline_instances_vs_oracle(cuda_device, width, nspokes):
    print(f"headline instances: W={width}, 32^3 x 32 coils, {nspokes}x64 golden-means", flush=True)
    rng = np.random.default_rng(40 + width)
    N, C = 32, 32
    coord = demo.golden_means_radial_3d(nspokes, 64)
    w = _ramp(coord)
    x = demo.complex_gaussian(rng, (N, N, N))
    maps = demo.gaussian_coil_maps(np.random.default_rng(2), C, (N, N, N))
    xc = (maps * x[None]).astype(np.complex64)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=width, oversamp=2.0)
    plan.to(cuda_device)
    what = _Call(plan, cuda_device, None, None, None).describe(C)
    assert f"gather=k_gather_col<{width}>" in what, what
    assert "spread=k_tile_max+k_spread_mfma<8,4,4>" in what, what
    P = R.Plan(coord, N, width=width, oversamp=2.0)

    y = pf.nufft(torch.from_numpy(xc).to(cuda_device), nufft_plan=plan, weight=torch.from_numpy(w),
                 device=cuda_device)
    yr = R.nufft(xc[:, None], P, weight=w)[:, 0].reshape((C,) + coord.shape[:-1])
    e_fwd = rel(y.cpu().numpy(), yr)

    k = demo.complex_gaussian(rng, (C,) + coord.shape[:-1])
    xa = pf.nufft_adj(torch.from_numpy(k).to(cuda_device), nufft_plan=plan, weight=torch.from_numpy(w),
                      device=cuda_device)
    xr = R.nufft_adj(k.reshape(C, 1, -1), P, weight=w)[:, 0]
    e_adj = rel(xa.cpu().numpy(), xr)
    print(f"  forward rel {e_fwd:.2e}, adjoint rel {e_adj:.2e}", flush=True)
    assert e_fwd < PARITY and e_adj < PARITY


def test_headline_full_size_properties(cuda_device):
    """128^3 x 32 coils, 8192 x 256 golden-means, sigma 2, |k|^2 weights."""
    print("headline full size: 128^3 x 32 coils, 8192x256, W=6", flush=True)
    N, C = 128, 32
    coord = demo.golden_means_radial_3d(8192, 256)
    w = torch.from_numpy(_ramp(coord)).to(cuda_device)
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0)
    plan.to(cuda_device)
    what = _Call(plan, cuda_device, None, None, None).describe(C)
    assert "gather=k_gather_col<6>" in what and "spread=k_tile_max+k_spread_mfma<8,4,4>" in what, what
    g = torch.Generator(device=cuda_device)
    g.manual_seed(7)
    x = torch.randn((C, N, N, N), dtype=torch.complex64, device=cuda_device, generator=g)
    y = torch.randn((C,) + coord.shape[:-1], dtype=torch.complex64, device=cuda_device, generator=g)

    # ortho adjointness with the weights on both sides: <W A x, y> = <x, A^H W y>
    Ax = pf.nufft(x, nufft_plan=plan, weight=w, norm="ortho", device=cuda_device)
    Ahy = pf.nufft_adj(y, nufft_plan=plan, weight=w, norm="ortho", device=cuda_device)
    lhs = torch.vdot(Ax.reshape(-1).cdouble(), y.reshape(-1).cdouble()).item()
    rhs = torch.vdot(x.reshape(-1).cdouble(), Ahy.reshape(-1).cdouble()).item()
    e_adj = abs(lhs - rhs) / abs(lhs)
    print(f"  adjointness {e_adj:.2e}", flush=True)
    assert e_adj < 1e-5

    # bitwise determinism of the adjoint (no float atomics)
    again = pf.nufft_adj(y, nufft_plan=plan, weight=w, norm="ortho", device=cuda_device)
    assert torch.equal(Ahy.view(torch.float32), again.view(torch.float32))
    del again, Ahy, Ax

    # forward vs the FP64 NUDFT on 1024 random samples of coil 0 (unweighted)
    y0 = pf.nufft(x[:1], nufft_plan=plan, device=cuda_device)
    sel = np.sort(np.random.default_rng(11).choice(coord.shape[0] * coord.shape[1], size=1024, replace=False))
    ct = plan.interpolator.coord.reshape(-1, 3).numpy()[sel]
    yt = U.nudft(x[0].cpu().numpy()[None], ct, [N, N, N], chunk=256)[0]
    e_nudft = rel(y0.reshape(-1)[torch.from_numpy(sel).to(cuda_device)].cpu().numpy(), yt)
    print(f"  W=6 forward vs FP64 NUDFT {e_nudft:.2e}", flush=True)
    assert e_nudft < 1e-4


def test_headline_staged_equals_fused(cuda_device):
    """The stage entry points (pad+FFT, gather, unsort; sort, spread,
    IFFT+crop) called one by one give the fused operators' result bit for
    bit.  The stand-alone spread computes the per-sample maxima of its
    matrix-core value scale itself (k_smax); the fused adjoint takes them from
    the sort."""
    from pyfourier_amd._nufft import stage_timer

    print("headline staged vs fused: 32^3 x 32 coils, 512x64, W=6", flush=True)
    rng = np.random.default_rng(9)
    N, C = 32, 32
    coord = demo.golden_means_radial_3d(512, 64)
    w = torch.from_numpy(_ramp(coord))
    plan = pf.plan_nufft(torch.from_numpy(coord), N, width=6, oversamp=2.0)
    plan.to(cuda_device)
    x = torch.from_numpy(demo.complex_gaussian(rng, (C, N, N, N))).to(cuda_device)
    k = torch.from_numpy(demo.complex_gaussian(rng, (C,) + coord.shape[:-1])).to(cuda_device)
    yf = pf.nufft(x, nufft_plan=plan, weight=w, device=cuda_device)
    xf = pf.nufft_adj(k, nufft_plan=plan, weight=w, device=cuda_device)
    with stage_timer():
        ys = pf.nufft(x, nufft_plan=plan, weight=w, device=cuda_device)
        xs = pf.nufft_adj(k, nufft_plan=plan, weight=w, device=cuda_device)
    torch.cuda.synchronize(cuda_device)
    assert torch.equal(yf.view(torch.float32), ys.view(torch.float32))
    assert torch.equal(xf.view(torch.float32), xs.view(torch.float32))
