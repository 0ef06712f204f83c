"""Golden fixtures (tests/golden/*.npz, written by make_golden.py).

CPU: the oracle still reproduces the committed outputs and the outputs sit
within the kernel-approximation bound of the FP64 NUDFT / exact field-map
model stored beside them.
GPU: the HIP engine reproduces the committed outputs at the 1e-5 bar.
"""

import os

import numpy as np
import pytest
import torch

from oracle import restate as R

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_oracle_reproduces_radial2d():
    d = load("radial2d")
    P = R.Plan(d["coord"], int(d["N"]), width=int(d["width"]), oversamp=float(d["oversamp"]))
    C = d["x"].shape[0]
    assert rel(R.nufft(d["x"][:, None], P)[:, 0], d["y"].reshape(C, -1)) < 1e-12
    assert rel(R.nufft_adj(d["k"].reshape(C, 1, -1), P)[:, 0], d["xadj"]) < 1e-12
    assert rel(d["y"], d["y_nudft"]) < 2e-3  # W=4, sigma=2 kernel error


def test_oracle_reproduces_radial3d():
    d = load("radial3d")
    P = R.Plan(d["coord"], int(d["N"]), width=int(d["width"]), oversamp=float(d["oversamp"]))
    assert rel(R.nufft(d["x"][None, None], P)[0, 0], d["y"].ravel()) < 1e-12
    assert rel(d["y"], d["y_nudft"]) < 2e-3  # W=5, sigma=1.5


def test_oracle_reproduces_subspace_and_b0():
    d = load("subspace")
    P = R.Plan(d["coord"], int(d["N"]), width=4, oversamp=2.0)
    assert rel(R.nufft(d["x"][None], P, basis=d["basis"])[0], d["y"].reshape(6, -1)) < 1e-12
    e = load("b0")
    P = R.Plan(e["coord"], int(e["N"]), width=6, oversamp=2.0)
    bt, ct = R.mri_exp_approx(e["zmap"], e["t"], L=int(e["L"]), nbins=40)
    assert rel(bt, e["bt"]) < 1e-10 and rel(ct, e["ct"]) < 1e-10
    assert rel(e["y"], e["y_exact"]) < 1e-2  # segmentation + kernel error


# ---- GPU -------------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_matches_fixtures(cuda_device):
    import pyfourier_amd as pf

    d = load("radial2d")
    plan = pf.plan_nufft(torch.from_numpy(d["coord"]), int(d["N"]), width=4, oversamp=2.0)
    y = pf.nufft(torch.from_numpy(d["x"]).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y.cpu().numpy(), d["y"]) < 1e-5
    xa = pf.nufft_adj(torch.from_numpy(d["k"]).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(xa.cpu().numpy(), d["xadj"]) < 1e-5

    d = load("radial3d")
    plan = pf.plan_nufft(torch.from_numpy(d["coord"]), int(d["N"]), width=5, oversamp=1.5)
    y = pf.nufft(torch.from_numpy(d["x"]).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y.cpu().numpy(), d["y"]) < 1e-5
    xa = pf.nufft_adj(torch.from_numpy(d["k"]).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(xa.cpu().numpy(), d["xadj"]) < 1e-5

    d = load("subspace")
    plan = pf.plan_nufft(torch.from_numpy(d["coord"]), int(d["N"]), width=4, oversamp=2.0)
    b = torch.from_numpy(d["basis"])
    y = pf.nufft(torch.from_numpy(d["x"]).to(cuda_device), nufft_plan=plan, basis=b, device=cuda_device)
    assert rel(y.cpu().numpy(), d["y"]) < 1e-5
    xa = pf.nufft_adj(torch.from_numpy(d["k"]).to(cuda_device), nufft_plan=plan, basis=b, device=cuda_device)
    assert rel(xa.cpu().numpy(), d["xadj"]) < 1e-5

    e = load("b0")
    plan = pf.plan_nufft(torch.from_numpy(e["coord"]), int(e["N"]), width=6, oversamp=2.0,
                         zmap=torch.from_numpy(e["zmap"]), T=torch.from_numpy(e["t"]), L=int(e["L"]), nbins=40)
    y = pf.nufft(torch.from_numpy(e["x"]).to(cuda_device), nufft_plan=plan, device=cuda_device)
    assert rel(y.cpu().numpy(), e["y"]) < 1e-5
