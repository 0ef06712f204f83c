"""Stage-level GPU parity through the C ABI: the fused deapodise/pad/FFT and
inverse FFT/crop/deapodise against the oracle's centred-FFT grid, for
1-D/2-D/3-D grids covering every radix of the fused FFT (8, 5, 4, 3, 2) and
the 256- and 512-point register column and row passes."""

import ctypes

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import restate as R
from pyfourier_amd import _lib, demo

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [
    # (shape, oversamp) -> grid sizes
    ((128,), 1.25),        # 160 = 8*5*4
    ((100,), 2.0),         # 200 = 8*5*5
    ((96,), 2.0),          # 192 = 8*8*3
    ((128,), 2.0),         # 256 = 8*8*4
    ((64, 128), 1.25),     # 80 x 160
    ((48, 40), 1.5),       # 72 x 60
    ((16, 20, 24), 2.0),   # 32 x 40 x 48
    ((128, 100), 2.0),     # 256 x 200: y is a 256-point column pass
    ((120, 96, 80), 2.0),  # 240 x 192 x 160 ... and 256-point z below
    ((128, 24, 20), 2.0),  # 256 x 48 x 40: z is a 256-point column pass
    ((12, 20, 128), 2.0),  # 24 x 40 x 256: x is a 256-point row pass
    ((10, 128), 2.0),      # 20 x 256
    ((256, 40), 2.0),      # 512 x 80: y is a 512-point column pass
    ((200, 25), 2.56),     # 512 x 64: 200 of the 512 y rows carry data
    ((256, 12, 10), 2.0),  # 512 x 24 x 20: z is a 512-point column pass
    ((12, 256, 10), 2.0),  # 24 x 512 x 20: y is a 512-point column pass
    ((256,), 2.0),         # 512: x is a 512-point row pass
    ((10, 256), 2.0),      # 20 x 512
    ((8, 200), 2.56),      # 16 x 512: 200 of the 512 x points carry data
    ((6, 8, 256), 2.0),    # 12 x 16 x 512
    ((64, 20), 2.0),       # 128 x 40: y is a 16 x 8 register column pass
    ((128, 20), 1.25),     # 160 x 25: 16 x 10
    ((100, 10), 1.6),      # 160 x 16: 100 of the 160 y rows carry data
    ((96, 24), 2.0),       # 192 x 48: 16 x 12
    ((192, 10, 8), 1.25),  # 240 x 13 x 10: z is a 16 x 15 register column pass
    ((150, 8), 1.6),       # 240 x 13
    ((256, 12), 1.25),     # 320 x 15: 16 x 20
    ((10, 192, 12), 2.0),  # 20 x 384 x 24: y is a 16 x 24 register column pass
    ((12, 64), 2.0),       # 24 x 128: x is a 16 x 8 register row pass
    ((10, 192), 1.25),     # 13 x 240: 16 x 15
    ((8, 256), 1.25),      # 10 x 320: 16 x 20
    ((6, 192), 2.0),       # 12 x 384: 16 x 24
    ((4, 6, 100), 1.6),    # 7 x 10 x 160: 100 of the 160 x points carry data
    ((384, 12), 1.25),     # 480 x 15: y is a 16 x 30 register column pass (30 = 6 x 5, 6 = 2 x 3)
    ((512, 8), 1.25),      # 640 x 10: 16 x 40
    ((10, 384), 1.25),     # 13 x 480: x is a 16 x 30 register row pass
    ((6, 512), 1.25),      # 8 x 640: 16 x 40
    ((384, 10, 8), 1.25),  # 480 x 13 x 10: z
    ((72, 20), 2.0),       # 144 x 40: 16 x 9 (Cooley-Tukey 3 x 3)
    ((144, 12), 2.0),      # 288 x 24: 16 x 18 (2 x 9)
    ((320, 10), 1.25),     # 400 x 13: 16 x 25 (Cooley-Tukey 5 x 5)
    ((288, 10), 2.0),      # 576 x 20: 16 x 36 (4 x 9)
    ((10, 72), 2.0),       # 20 x 144: the same as row passes
    ((8, 144), 2.0),       # 16 x 288
    ((6, 320), 1.25),      # 8 x 400
    ((6, 288), 2.0),       # 12 x 576
    ((64, 20), 1.25),      # 80 x 25: 16 x 5
    ((48, 10, 8), 2.0),    # 96 x 20 x 16: z is 16 x 6
    ((10, 64), 1.25),      # 13 x 80: rows
    ((12, 48), 2.0),       # 24 x 96
    ((96, 20), 1.25),      # 120 x 25: 8 threads x 15 values per column
    ((160, 12), 1.25),     # 200 x 15: 8 x 25
    ((180, 10), 2.0),      # 360 x 20: 8 x 45
    ((96, 10, 8), 1.25),   # 120 x 13 x 10: z
    ((10, 96), 1.25),      # 13 x 120: rows
    ((8, 160), 1.25),      # 10 x 200
    ((6, 180), 2.0),       # 12 x 360
    ((56, 20), 2.0),       # 112 x 40: 16 x 7 (7-point DFT)
    ((112, 12), 2.0),      # 224 x 24: 16 x 14 (2 x 7)
    ((168, 10), 2.0),      # 336 x 20: 16 x 21 (3 x 7)
    ((224, 8), 2.0),       # 448 x 16: 16 x 28 (4 x 7)
    ((224, 10), 1.25),     # 280 x 13: 8 x 35 (5 x 7)
    ((10, 112), 2.0),      # 20 x 224: rows
    ((6, 224), 1.25),      # 8 x 280
    ((6, 224), 2.0),       # 12 x 448
    ((32, 20), 1.25),      # 40 x 25: 8 x 5
    ((28, 10, 8), 2.0),    # 56 x 20 x 16: 8 x 7
    ((36, 20), 2.0),       # 72 x 40: 8 x 9
    ((84, 12), 2.0),       # 168 x 24: 8 x 21
    ((10, 32), 1.25),      # 13 x 40: rows
    ((12, 36), 2.0),       # 24 x 72
    ((6, 84), 2.0),        # 12 x 168
]
# (shape, oversamp, coils): 3 coils, plus 512-point row passes with 1 or 2
# values per grid point (second-stage lanes m1-fastest) and with 4
PARAMS = [(s, o, 3) for s, o in CASES] + [
    ((256,), 2.0, 2),
    ((10, 256), 2.0, 1),
    ((8, 200), 2.56, 2),
    ((6, 8, 256), 2.0, 2),
    ((6, 8, 256), 2.0, 4),
    ((6, 8, 128), 1.25, 2),  # 160-point row pass storing m1-fastest (coil pairs)
    ((6, 8, 192), 2.0, 16),
    ((6, 8, 96), 1.25, 2),   # 120-point row pass (8 threads per row) storing m1-fastest
    ((6, 8, 160), 1.25, 12),
]


def _desc(C):
    d = _lib.BatchDesc()
    d.batch, d.nseg, d.ncoeff, d.nt = C, 1, 0, 1
    return d


@pytest.mark.parametrize("shape,oversamp,C", PARAMS)
def test_pad_fft_stage(cuda_device, shape, oversamp, C):
    _check_stages(cuda_device, shape, oversamp, C)


def test_pad_fft_stage_wide_512_columns(cuda_device):
    """A 512-point z pass over 2^18 columns (16 coils x 128 x 128), where the
    16-thread-per-column kernel takes over from the 32-thread one."""
    _check_stages(cuda_device, (256, 64, 64), 2.0, 16)


def _check_stages(cuda_device, shape, oversamp, C):
    rng = np.random.default_rng(0)
    d = len(shape)
    coord = rng.uniform(-np.pi, np.pi, size=(256, d)).astype(np.float32)
    coord[0] = 0
    coord[0, 0] = -np.pi
    plan = pf.plan_nufft(torch.from_numpy(coord), shape, width=4, oversamp=oversamp)
    plan.to(cuda_device)
    P = R.Plan(coord, shape, width=4, oversamp=oversamp)
    nat = plan.interpolator.native(cuda_device)
    x = demo.complex_gaussian(rng, (C,) + tuple(shape))
    desc = _desc(C)
    L = _lib.lib()
    off = [ctypes.c_size_t() for _ in range(4)]
    _lib.check(L.pf_workspace_offsets(nat.handle, ctypes.byref(desc), *[ctypes.byref(o) for o in off]))
    nbytes = ctypes.c_size_t()
    _lib.check(L.pf_workspace_size(nat.handle, ctypes.byref(desc), 1, ctypes.byref(nbytes)))
    ws = torch.zeros(nbytes.value, dtype=torch.uint8, device=cuda_device)
    grid = ws[off[0].value: off[1].value].view(torch.complex64)
    fws = ws[off[2].value: off[3].value]
    xd = torch.from_numpy(x).to(cuda_device)
    st = _lib.stream_of(cuda_device)
    _lib.check(L.pf_pad_fft(nat.handle, ctypes.byref(desc), _lib.ptr(xd), _lib.ptr(grid), _lib.ptr(fws), st))
    torch.cuda.synchronize()
    G = P.os_shape
    # grid layout (C / Cv, *G, Cv): Cv is the largest divisor of C up to 64,
    # or 2 (coil pairs) for a 3-D plan with one kernel width and an even C
    Cv = 2 if d == 3 and C % 2 == 0 else max(v for v in range(1, min(C, 64) + 1) if C % v == 0)
    g = grid[: C * int(np.prod(G))].cpu().numpy().reshape((C // Cv,) + tuple(G) + (Cv,))
    g = np.moveaxis(g, -1, 1).reshape((C,) + tuple(G))
    ref = R._to_grid(x[:, None], P, None)[:, 0]
    ref = np.fft.ifftshift(ref, axes=tuple(range(-d, 0)))  # stored (unshifted) order
    err = rel(g, ref)
    assert err < 2e-6, f"pad_fft {shape} sigma={oversamp} grid={G}: rel err {err:.3e}"
    # adjoint stage: inverse FFT, crop, deapodise (factor 1), from the
    # reference grid in the adjoint's layout (Cv the largest divisor of C up to 64)
    Cva = max(v for v in range(1, min(C, 64) + 1) if C % v == 0)
    ga = np.moveaxis(ref.reshape((C // Cva, Cva) + tuple(G)), 1, -1).astype(np.complex64)
    grid[: C * int(np.prod(G))].copy_(torch.from_numpy(ga.ravel()))
    img = torch.empty((C,) + tuple(shape), dtype=torch.complex64, device=cuda_device)
    _lib.check(L.pf_ifft_crop(nat.handle, ctypes.byref(desc), _lib.ptr(grid), _lib.ptr(img), 1.0,
                              _lib.ptr(fws), st))
    torch.cuda.synchronize()
    back = R._from_grid(np.fft.fftshift(ref, axes=tuple(range(-d, 0)))[:, None], P, None)[:, 0]
    back = back * np.prod(G)  # _from_grid uses the 1/Ghat inverse; the stage is unnormalised
    err = rel(img.cpu().numpy(), back)
    assert err < 2e-6, f"ifft_crop {shape} sigma={oversamp}: rel err {err:.3e}"
