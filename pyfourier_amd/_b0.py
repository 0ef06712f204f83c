"""Time-segmented field-map model (host side, run once per plan).

Approximates exp(-z(r) t) by sum_l B[l, t] Ct[l, r] with the histogram-
weighted least-squares interpolator of MIRT mri_exp_approx ("hist,time,unif"),
following src/pyfourier/_subroutines/_b0.py:10-135 with defect A.15 fixed:

* bin centres are (e[i] + e[i+1]) / 2 (the reference uses e[1:] - e[1]/2);
* the temporal basis fits exp(-z t), the same sign as Ct = exp(-tl z), so
  that sum_l B[l,t] Ct[l,r] approximates exp(-z(r) t) (the reference fits
  exp(+z t));
* the module inputs (backend, device) come from the field map itself.

z = 2 pi (R2* + i df) for a complex map, z = 2 pi i df for a real one.
Segment times tl = l/(L-1) * t_end (src/pyfourier/_subroutines/_b0.py:113).
"""

from __future__ import annotations

import math

import torch


def mri_exp_approx(zmap, t, L: int = 6, nbins=(40, 40), mask=None):
    """Return (B (L, nt) complex64, Ct (L, *zmap.shape) complex64) on zmap's device."""
    zmap = torch.as_tensor(zmap)
    device = zmap.device
    t = torch.as_tensor(t).to(torch.float64).reshape(-1).cpu()
    zc = zmap.detach().cpu()
    if mask is None:
        mask = torch.ones(zc.shape, dtype=torch.bool)
    else:
        mask = torch.as_tensor(mask, dtype=torch.bool).cpu()
    if not torch.is_complex(zc):
        z = 2j * math.pi * zc.to(torch.float64)
        nb = nbins[0] if isinstance(nbins, (list, tuple)) else int(nbins)
        vals = z.imag[mask].reshape(-1)
        lo, hi = float(vals.min()), float(vals.max())
        if hi <= lo:
            hi = lo + 1e-9
        hk = torch.histc(vals, bins=nb, min=lo, max=hi).to(torch.float64)
        e = torch.linspace(lo, hi, nb + 1, dtype=torch.float64)
        zk = 1j * (e[:-1] + e[1:]) / 2
    else:
        z = 2 * math.pi * zc.to(torch.complex128)
        nb = tuple(nbins) if isinstance(nbins, (list, tuple)) else (int(nbins), 40)
        vi = z.imag[mask].reshape(-1)
        vr = z.real[mask].reshape(-1)
        hk, edges = torch.histogramdd(torch.stack([vi, vr], dim=1), bins=list(nb))
        ei, er = edges
        ci = (ei[:-1] + ei[1:]) / 2
        cr = (er[:-1] + er[1:]) / 2
        zk = 1j * ci.to(torch.float64)[:, None] + cr.to(torch.float64)[None, :]
        hk = hk.to(torch.float64)
    hk = hk.reshape(-1).to(torch.complex128)
    zk = zk.reshape(-1).to(torch.complex128)
    tl = torch.linspace(0, L, L, dtype=torch.float64) / L * t[-1]
    tlc = tl.to(torch.complex128)
    ch = torch.exp(-tlc[:, None] * zk[None, :])  # (L, nk)
    w = hk.sqrt()
    p = torch.linalg.pinv(w[:, None] * ch.T) * w[None, :]  # (L, nk)
    b = p @ torch.exp(-zk[:, None] * t.to(torch.complex128)[None, :])  # (L, nt)
    ct = torch.exp(-tlc.reshape((-1,) + (1,) * z.ndim) * z[None])
    ct = torch.nan_to_num(ct, nan=0.0, posinf=0.0, neginf=0.0)
    return b.to(torch.complex64).to(device), ct.to(torch.complex64).to(device)
