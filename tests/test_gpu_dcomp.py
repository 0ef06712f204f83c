"""GPU parity of pipe_menon_dcf (graph-replayed iteration) against the CPU
restatement of src/pyfourier/_dcomp/pipemenon.py.

The iteration w <- w / |G G^H w| with the full NUFFT amplifies rounding
wherever G G^H w comes close to zero: on the CPU restatement alone a 1e-7
relative perturbation of the start grows to ~2e-1 (3-D) after three
iterations.  Parity is therefore asserted for the first two iterations; the
Cartesian fixed point pins any number of iterations exactly, and the graph
replay must equal the eager loop bit for bit.
"""

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


@pytest.mark.parametrize("niter", [1, 2])
def test_pipe_menon_2d_parity(cuda_device, niter):
    coord = demo.golden_angle_radial_2d(40, 96)
    w = pf.pipe_menon_dcf(torch.from_numpy(coord), 48, niter=niter, device=cuda_device)
    assert w.shape == coord.shape[:-1] and w.dtype == torch.float32 and w.device.type == "cpu"
    wr = R.pipe_menon_dcf(coord, 48, niter=niter)
    assert rel(w.numpy(), wr) < 1e-4


def test_pipe_menon_3d_graph_equals_eager_and_numpy(cuda_device):
    coord = demo.golden_means_radial_3d(150, 32)
    wg = pf.pipe_menon_dcf(coord, 16, niter=5, device=cuda_device)
    we = pf.pipe_menon_dcf(coord, 16, niter=5, device=cuda_device, graph=False)
    assert isinstance(wg, np.ndarray) and wg.shape == coord.shape[:-1]
    assert np.array_equal(wg, we)
    # one iteration: in 3-D the second already amplifies float32 rounding
    # ~1e3-fold; even w = 1 / |G G^H 1| carries the relative error of G G^H 1
    # where it is small through cancellation, hence 5e-4
    w1 = pf.pipe_menon_dcf(coord, 16, niter=1, device=cuda_device)
    assert rel(w1, R.pipe_menon_dcf(coord, 16, niter=1)) < 5e-4


def test_pipe_menon_cartesian_fixed_point(cuda_device):
    g = np.arange(32) - 16
    cc = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(32, 32, 2).astype(np.float32)
    w = pf.pipe_menon_dcf(cc, 32, niter=12, device=cuda_device)
    assert np.max(np.abs(w - 1.0)) < 1e-5
