"""GPU parity of pipe_menon_dcf (graph-replayed iteration) against the CPU
restatement of src/pyfourier/_dcomp/pipemenon.py."""

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


@pytest.mark.parametrize("niter", [1, 4])
def test_pipe_menon_2d_parity(cuda_device, niter):
    coord = demo.golden_angle_radial_2d(40, 96)
    w = pf.pipe_menon_dcf(torch.from_numpy(coord), 48, niter=niter, device=cuda_device)
    assert w.shape == coord.shape[:-1] and w.dtype == torch.float32 and w.device.type == "cpu"
    wr = R.pipe_menon_dcf(coord, 48, niter=niter)
    assert rel(w.numpy(), wr) < 1e-4 * niter


def test_pipe_menon_graph_equals_eager_and_numpy(cuda_device):
    coord = demo.golden_means_radial_3d(150, 32)
    wg = pf.pipe_menon_dcf(coord, 16, niter=3, device=cuda_device)
    we = pf.pipe_menon_dcf(coord, 16, niter=3, device=cuda_device, graph=False)
    assert isinstance(wg, np.ndarray) and wg.shape == coord.shape[:-1]
    assert np.array_equal(wg, we)
    wr = R.pipe_menon_dcf(coord, 16, niter=3)
    assert rel(wg, wr) < 3e-4
