"""FP64 exact non-uniform DFT: the ground truth the NUFFT approximates.

TEST INFRASTRUCTURE ONLY (see oracle/restate.py for who may import it).

Convention (the unit-gain limit of the reference pipeline,
src/pyfourier/_nufft/_nufft.py:218-287 with src/pyfourier/_subroutines/_fft.py:72-79):

    y_j = sum_r x_r exp(-2 pi i sum_a k_{j,a} r_a / N_a),   r_a = p_a - N_a//2

with k in grid units [-N/2, N/2) and image axes ordered (..., z, y, x) while
the coordinate's last axis is ordered (x, y, z).  The adjoint uses exp(+...).
Evaluation is separable: one phase matrix per axis, contracted with BLAS, in
chunks of samples so memory stays bounded.
"""

from __future__ import annotations

import math

import numpy as np


def _phase(k_axis, n, sign):
    r = np.arange(n, dtype=np.float64) - n // 2
    return np.exp(sign * 2j * math.pi * np.outer(np.asarray(k_axis, dtype=np.float64), r) / n)


def nudft(x, coord, shape, chunk=2048):
    """x (B, *N) complex, coord (M, d) grid units (x, y, z) -> y (B, M)."""
    x = np.asarray(x, dtype=np.complex128)
    coord = np.asarray(coord, dtype=np.float64).reshape(-1, len(shape))
    d = len(shape)
    B = x.shape[0]
    M = coord.shape[0]
    y = np.zeros((B, M), dtype=np.complex128)
    for s in range(0, M, chunk):
        c = coord[s : s + chunk]
        # phase per axis in image order (z, y, x): image axis a pairs with coord[:, d-1-a]
        ph = [_phase(c[:, d - 1 - a], shape[a], -1.0) for a in range(d)]  # (m, N_a)
        if d == 1:
            y[:, s : s + chunk] = x @ ph[0].T
        elif d == 2:
            t = np.einsum("byx,mx->bym", x, ph[1], optimize=True)
            y[:, s : s + chunk] = np.einsum("bym,my->bm", t, ph[0], optimize=True)
        else:
            t = np.einsum("bzyx,mx->bzym", x, ph[2], optimize=True)
            t = np.einsum("bzym,my->bzm", t, ph[1], optimize=True)
            y[:, s : s + chunk] = np.einsum("bzm,mz->bm", t, ph[0], optimize=True)
    return y


def nudft_adj(y, coord, shape, chunk=2048):
    """y (B, M) -> x (B, *N): x_r = sum_j y_j exp(+2 pi i k_j . r / N)."""
    y = np.asarray(y, dtype=np.complex128)
    coord = np.asarray(coord, dtype=np.float64).reshape(-1, len(shape))
    d = len(shape)
    B = y.shape[0]
    x = np.zeros((B,) + tuple(shape), dtype=np.complex128)
    for s in range(0, coord.shape[0], chunk):
        c = coord[s : s + chunk]
        ys = y[:, s : s + chunk]
        ph = [_phase(c[:, d - 1 - a], shape[a], 1.0) for a in range(d)]
        if d == 1:
            x += ys @ ph[0]
        elif d == 2:
            x += np.einsum("bm,my,mx->byx", ys, ph[0], ph[1], optimize=True)
        else:
            t = np.einsum("bm,mz->bzm", ys, ph[0], optimize=True)
            t = np.einsum("bzm,my->bzym", t, ph[1], optimize=True)
            x += np.einsum("bzym,mx->bzyx", t, ph[2], optimize=True)
    return x


def nudft_b0(x, coord, shape, zmap, t, chunk=256):
    """Exact field-map model y_j = sum_r x_r exp(-z_r t_j) exp(-2 pi i k_j.r/N).

    x (B, *N); coord (M, d) with M = nviews*nsamples ordered (view, sample);
    zmap real (Hz, z = 2 pi i f) or complex (R2* + i f, z = 2 pi zmap);
    t (nsamples,) seconds.  Direct evaluation: small problems only.
    """
    x = np.asarray(x, dtype=np.complex128)
    d = len(shape)
    coord = np.asarray(coord, dtype=np.float64).reshape(-1, d)
    zmap = np.asarray(zmap)
    z = (2j * math.pi * zmap) if np.isrealobj(zmap) else (2 * math.pi * zmap.astype(np.complex128))
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    tj = np.tile(t, coord.shape[0] // t.size)
    grids = np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij")
    r = np.stack([g.ravel() for g in grids], axis=1).astype(np.float64)  # (P, d) in (z,y,x)
    nvec = np.asarray(shape, dtype=np.float64)
    xf = x.reshape(x.shape[0], -1)
    zf = z.ravel()
    y = np.zeros((x.shape[0], coord.shape[0]), dtype=np.complex128)
    for s in range(0, coord.shape[0], chunk):
        c = coord[s : s + chunk][:, ::-1]  # to (z, y, x)
        ph = np.exp(-2j * math.pi * (c / nvec) @ r.T - tj[s : s + chunk, None] * zf[None, :])
        y[:, s : s + chunk] = xf @ ph.T
    return y


def psf_exact(coord, shape, weight=None):
    """PSF(dr) = sum_j w_j exp(+2 pi i k_j . dr / N) on dr in [-N, N)^d.

    Returned on a 2N grid with dr = p - N, i.e. the adjoint NUDFT of w on the
    2N grid at coordinates 2k.
    """
    coord = np.asarray(coord, dtype=np.float64).reshape(-1, len(shape))
    w = np.ones(coord.shape[0]) if weight is None else np.asarray(weight, dtype=np.float64).reshape(-1)
    big = [2 * n for n in shape]
    return nudft_adj(w[None].astype(np.complex128), 2.0 * coord, big)[0]
