"""Seeded synthetic inputs for tests and benchmarks (counterpart of the
reference's src/pyfourier/demo/ generators).

All trajectories are returned in radians, k in [-pi, pi), shaped
(..., ndim) with the last axis ordered (x, y, z); plan_nufft maps them to
grid units.
"""

from __future__ import annotations

import math

import numpy as np

GOLDEN = (1 + 5 ** 0.5) / 2


def complex_gaussian(rng: np.random.Generator, shape) -> np.ndarray:
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2)).astype(np.complex64)


def golden_angle_radial_2d(nspokes: int, nsamples: int) -> np.ndarray:
    """(nspokes, nsamples, 2); spoke n at angle n*pi/phi (111.246 deg)."""
    ang = np.arange(nspokes) * math.pi / GOLDEN
    k = np.linspace(-math.pi, math.pi, nsamples, endpoint=False)
    c = np.stack([np.cos(ang)[:, None] * k[None], np.sin(ang)[:, None] * k[None]], axis=-1)
    return c.astype(np.float32)


def golden_means_radial_3d(nspokes: int, nsamples: int) -> np.ndarray:
    """(nspokes, nsamples, 3) golden-means 3D radial (Chan et al. 2009):
    cos(polar) = n*0.4656 mod 1, azimuth = 2 pi (n*0.6823 mod 1)."""
    n = np.arange(nspokes)
    z = np.mod(n * 0.46557123, 1.0)
    az = 2 * math.pi * np.mod(n * 0.68232780, 1.0)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(az), r * np.sin(az), z], axis=-1)
    k = np.linspace(-math.pi, math.pi, nsamples, endpoint=False)
    return (d[:, None, :] * k[None, :, None]).astype(np.float32)


def variable_density_spiral_2d(ninterleaves: int, nsamples: int, npix: int, alpha: float = 1.5) -> np.ndarray:
    """(ninterleaves, nsamples, 2): k(t) = k_max tau^alpha exp(i 2 pi n_turns tau^alpha),
    interleaves rotated by 2 pi i / ninterleaves; k_max = pi."""
    tau = np.linspace(0, 1, nsamples, endpoint=False)
    turns = max(1.0, npix / (2.0 * ninterleaves))
    r = math.pi * tau ** alpha
    th = 2 * math.pi * turns * tau ** alpha
    rot = 2 * math.pi * np.arange(ninterleaves) / ninterleaves
    ang = th[None, :] + rot[:, None]
    c = np.stack([r[None] * np.cos(ang), r[None] * np.sin(ang)], axis=-1)
    # pin the global radius to exactly pi so normalisation maps it to N/2
    c[0, -1] = [math.pi * math.cos(ang[0, -1]), math.pi * math.sin(ang[0, -1])]
    return c.astype(np.float32)


def spiral_density(coord: np.ndarray) -> np.ndarray:
    """Analytic density weights |k| |dk/dt| (normalised to max 1) for spirals."""
    k = coord[..., 0] + 1j * coord[..., 1]
    dk = np.gradient(k, axis=-1)
    w = np.abs(k) * np.abs(dk)
    return (w / w.max()).astype(np.float32)


def gaussian_coil_maps(rng: np.random.Generator, ncoils: int, shape) -> np.ndarray:
    """Smooth complex Gaussian sensitivity maps, RSS-normalised."""
    d = len(shape)
    grids = np.meshgrid(*[np.linspace(-0.5, 0.5, n, endpoint=False) for n in shape], indexing="ij")
    maps = np.zeros((ncoils,) + tuple(shape), dtype=np.complex128)
    for c in range(ncoils):
        ctr = rng.uniform(-0.5, 0.5, size=d)
        r2 = sum((g - c0) ** 2 for g, c0 in zip(grids, ctr))
        maps[c] = np.exp(-r2 / (2 * 0.3 ** 2)) * np.exp(1j * rng.uniform(0, 2 * math.pi))
    rss = np.sqrt((np.abs(maps) ** 2).sum(axis=0))
    return (maps / rss).astype(np.complex64)


def smooth_field_map(rng: np.random.Generator, shape, max_hz: float = 100.0) -> np.ndarray:
    """Smooth Gaussian-bump field map scaled to +-max_hz (real, Hz)."""
    grids = np.meshgrid(*[np.linspace(-0.5, 0.5, n, endpoint=False) for n in shape], indexing="ij")
    f = np.zeros(shape)
    for _ in range(3):
        ctr = rng.uniform(-0.3, 0.3, size=len(shape))
        s = rng.uniform(0.1, 0.3)
        f += rng.choice([-1, 1]) * np.exp(-sum((g - c0) ** 2 for g, c0 in zip(grids, ctr)) / (2 * s * s))
    f = f / np.abs(f).max() * max_hz
    return f.astype(np.float32)


def t2_subspace_basis(rng: np.random.Generator, ncontrasts: int, ncoeff: int, nsig: int = 1000) -> np.ndarray:
    """(ncontrasts, ncoeff) orthonormal basis from the SVD of exp(-t/T2),
    t = linspace(1, 300, ncontrasts) ms, T2 ~ U[20, 300] ms (convention of
    the reference's demo/_tse.py:28-40)."""
    t = np.linspace(1, 300, ncontrasts)
    t2 = rng.uniform(20, 300, size=nsig)
    sig = np.exp(-t[None, :] / t2[:, None])
    _, _, vh = np.linalg.svd(sig, full_matrices=False)
    return vh[:ncoeff].T.astype(np.complex64)
