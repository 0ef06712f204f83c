"""GPU Voronoi density compensation (pf_voronoi_area_2d + pyfourier_amd.voronoi)
against the oracle (scipy's Qhull tessellation with the reference's
post-processing, oracle/voronoi.py; src/pyfourier/_dcomp/voronoi.py:16-123).

Cell areas are compared in float64 (relative 1e-9; open cells must match
exactly), the weights in float32 (relative 1e-5).
"""

import math

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import voronoi as V
from pyfourier_amd import demo

pytestmark = pytest.mark.gpu


def _check_areas(cuda_device, pts, label):
    a = pf.voronoi.voronoi_areas(torch.from_numpy(pts), cuda_device).cpu().numpy()
    b = V.cell_areas(pts)
    open_a, open_b = a == 0, b == 0
    rel = np.abs(a - b)[~open_b] / b[~open_b]
    print(f"  {label}: {len(pts)} points, open {int(open_b.sum())}, max rel {rel.max():.2e}", flush=True)
    assert np.array_equal(open_a, open_b)
    assert rel.max() < 1e-9


def test_voronoi_areas_random_and_lattices(cuda_device):
    rng = np.random.default_rng(3)
    _check_areas(cuda_device, rng.uniform(-1, 1, (5000, 2)), "uniform random")
    _check_areas(cuda_device, rng.standard_normal((3000, 2)) * [1.0, 0.1], "anisotropic gaussian")
    g = np.stack(np.meshgrid(np.arange(40.0), np.arange(31.0)), -1).reshape(-1, 2) * 0.1
    _check_areas(cuda_device, g, "square grid (cocircular vertices)")
    hexa = np.asarray([((i + 0.5 * (j % 2)) * 0.2, j * 0.2 * math.sqrt(3) / 2) for j in range(30) for i in range(30)])
    _check_areas(cuda_device, hexa, "hexagonal lattice")


def test_voronoi_areas_radial_centre_cell(cuda_device):
    """64 full spokes: the k-space centre's cell has 128 edges, more than the
    first pass holds, and is measured again with device-memory polygons."""
    spokes = demo.golden_angle_radial_2d(64, 64).reshape(-1, 2).astype(np.float32).astype(np.float64)
    u = np.unique(spokes, axis=0)
    _check_areas(cuda_device, u, "radial 64 x 64")


def test_voronoi_weights_spiral_full_disk(cuda_device):
    """The reference's own use (demo/_spiral.py:72): one interleaf, nshots
    copies rotated into a full disk, fullspoke=True."""
    k = demo.variable_density_spiral_2d(1, 1024, 128)  # (1, npts, 2)
    w = pf.voronoi.voronoi(k, nshots=16, fullspoke=True, device=cuda_device)
    wr = V.voronoi(k, 16, True)
    assert isinstance(w, np.ndarray) and w.dtype == np.float32 and w.shape == wr.shape
    e = float(np.max(np.abs(w - wr)) / np.max(np.abs(wr)))
    print(f"  spiral full disk: max rel {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("fix_edge", [False, True])
def test_voronoi_weights_radial_duplicates(cuda_device, fix_edge):
    """A full radial spoke through k = 0 rotated 48 times: the centre is
    repeated 48 times (split evenly), the spoke ends are open cells (0,
    extrapolated from the outer ring with fix_edge)."""
    spoke = np.stack([np.linspace(-math.pi, math.pi, 129), np.zeros(129)], -1)[None].astype(np.float32)
    w = pf.voronoi.voronoi(torch.from_numpy(spoke).to(cuda_device), nshots=48, fullspoke=True, fix_edge=fix_edge)
    wr = V.voronoi(spoke, 48, True, fix_edge=fix_edge)
    assert isinstance(w, torch.Tensor) and w.device.type == "cuda"
    e = float(np.max(np.abs(w.cpu().numpy() - wr)) / np.max(np.abs(wr)))
    print(f"  radial full disk, fix_edge={fix_edge}: max rel {e:.2e}", flush=True)
    assert e < 1e-5


def test_voronoi_weights_complex_and_float64(cuda_device):
    rng = np.random.default_rng(5)
    p = rng.uniform(-1, 1, (2, 300, 2))
    w = pf.voronoi.voronoi(p, nshots=1, fullspoke=True, device=cuda_device)  # float64 kept
    wr = V.voronoi(p, 1, True)
    assert w.shape == (2, 300) and np.max(np.abs(w - wr)) / np.max(wr) < 1e-5
    pc = (p[..., 0] + 1j * p[..., 1]).astype(np.complex64)
    wc = pf.voronoi.voronoi(pc, nshots=1, fullspoke=True, device=cuda_device)
    wcr = V.voronoi(pc, 1, True)
    assert np.max(np.abs(wc - wcr)) / np.max(wcr) < 1e-5


def test_voronoi_rejects_non_finite(cuda_device):
    p = np.zeros((10, 2))
    p[:, 0] = np.arange(10)
    p[3, 1] = np.nan
    with pytest.raises(ValueError):
        pf.voronoi.voronoi_areas(p, cuda_device)
