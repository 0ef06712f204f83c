"""Write the golden fixtures of the NUFFT path (small .npz files).

Inputs are seeded synthetic data (pyfourier_amd.demo); expected outputs come
from the oracle's CPU restatement (oracle/restate.py) and the FP64 NUDFT
(oracle/nudft.py).  The reference implementation itself cannot run (numba is
absent, and its import fails without an NVIDIA driver), so no fixture comes
from it.  Usage: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import nudft as U  # noqa: E402
from oracle import restate as R  # noqa: E402
from pyfourier_amd import demo  # noqa: E402


def radial2d():
    rng = np.random.default_rng(100)
    N, C = 32, 2
    coord = demo.golden_angle_radial_2d(16, 64)
    x = demo.complex_gaussian(rng, (C, N, N))
    k = demo.complex_gaussian(rng, (C, 16, 64))
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    return dict(
        coord=coord, x=x, k=k, N=N, width=4, oversamp=2.0,
        y=R.nufft(x[:, None], P)[:, 0].reshape(C, 16, 64),
        xadj=R.nufft_adj(k.reshape(C, 1, -1), P)[:, 0],
        y_nudft=U.nudft(x, P.coord.reshape(-1, 2), [N, N]).reshape(C, 16, 64),
    )


def radial3d():
    rng = np.random.default_rng(101)
    N = 12
    coord = demo.golden_means_radial_3d(48, 24)
    x = demo.complex_gaussian(rng, (N, N, N))
    k = demo.complex_gaussian(rng, (48, 24))
    P = R.Plan(coord, N, width=5, oversamp=1.5)
    return dict(
        coord=coord, x=x, k=k, N=N, width=5, oversamp=1.5,
        y=R.nufft(x[None, None], P)[0, 0].reshape(48, 24),
        xadj=R.nufft_adj(k.reshape(1, 1, -1), P)[0, 0],
        y_nudft=U.nudft(x[None], P.coord.reshape(-1, 3), [N, N, N])[0].reshape(48, 24),
    )


def subspace():
    rng = np.random.default_rng(102)
    F, K, N = 6, 3, 16
    base = demo.golden_angle_radial_2d(2, 32)
    coord = np.stack([base @ np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]]).T
                      for a in 0.9 * np.arange(F)]).astype(np.float32)
    basis = demo.t2_subspace_basis(rng, F, K)
    x = demo.complex_gaussian(rng, (K, N, N))
    k = demo.complex_gaussian(rng, (F, 2, 32))
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    return dict(
        coord=coord, basis=basis, x=x, k=k, N=N, width=4, oversamp=2.0,
        y=R.nufft(x[None], P, basis=basis)[0].reshape(F, 2, 32),
        xadj=R.nufft_adj(k.reshape(1, F, -1), P, basis=basis)[0],
    )


def b0():
    rng = np.random.default_rng(103)
    N = 16
    coord = demo.variable_density_spiral_2d(2, 128, N)
    zmap = demo.smooth_field_map(rng, (N, N), 100.0)
    t = np.linspace(0, 8e-3, 128).astype(np.float32)
    bt, ct = R.mri_exp_approx(zmap, t, L=6, nbins=40)
    x = demo.complex_gaussian(rng, (N, N))
    P = R.Plan(coord, N, width=6, oversamp=2.0)
    return dict(
        coord=coord, zmap=zmap, t=t, x=x, N=N, width=6, oversamp=2.0, L=6,
        bt=bt, ct=ct,
        y=R.nufft(x[None, None], P, b0=(bt, ct))[0, 0].reshape(2, 128),
        y_exact=U.nudft_b0(x[None], P.coord.reshape(-1, 2), (N, N), zmap, t)[0].reshape(2, 128),
    )


CASES = {"radial2d": radial2d, "radial3d": radial3d, "subspace": subspace, "b0": b0}

if __name__ == "__main__":
    for name, fn in CASES.items():
        d = fn()
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
        print("wrote", name)
