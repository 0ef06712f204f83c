"""CPU restatement of pyfourier's Voronoi density compensation (numpy + scipy).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker, never by the
product.  It follows src/pyfourier/_dcomp/voronoi.py (paths relative to the
reference repository root):

* tessellation   scipy.spatial.Voronoi (Qhull), the reference's own
                 dependency (voronoi.py:11,330-343; scipy 1.15.3 here);
* cell areas     the polygon formula over Qhull's region vertices for every
                 region without -1 (voronoi.py:262-286);
* spike clipping the n-th largest area, n = round((1 - (1 - ratio) * 0.98) M)
                 (voronoi.py:236-248, _select_nth :360-366);
* edge fix       cubic fit of area against |k|^2 over |k|^2 > 0.6 max, for
                 the open cells (voronoi.py:250-258);
* repeated pts   each copy gets its location's area / copies (the intent of
                 voronoi.py:20; the reference's copy lookup :199-217 mixes
                 unique ids with point indices);
* full disk      the first shot rotated about z by the shot angles
                 (voronoi.py:66-72, :374-429);
* normalisation  _normalize_weight (voronoi.py:186-194), then the k-space
                 centre scaled to 1 / nshots (voronoi.py:115-119).

2-D only, like the product.
"""

import math

import numpy as np
from scipy.spatial import Voronoi


def cell_areas(points):
    """Area of each distinct 2-D point's bounded Voronoi cell, 0 if open."""
    p = np.asarray(points, dtype=np.float64)
    vor = Voronoi(p)
    out = np.zeros(len(p))
    for n in range(len(p)):
        reg = vor.regions[vor.point_region[n]]
        if len(reg) == 0 or -1 in reg:
            continue
        v = vor.vertices[reg]
        a = 0.0
        for i in range(1, len(v) - 1):
            a += v[i, 0] * (v[i + 1, 1] - v[i - 1, 1])
        a += v[-1, 0] * (v[0, 1] - v[-2, 1])
        a += v[0, 0] * (v[1, 1] - v[-1, 1])
        out[n] = abs(a) / 2.0
    return out


def _clip_and_fix(wi, pts, fix_edge, ratio):
    fufa = 0.98
    while True:
        r = 1 - (1 - ratio) * fufa
        thld = np.sort(wi)[::-1][min(round(r * wi.size), wi.size - 1)]
        if np.isfinite(thld):
            break
        fufa **= 2
    wi = np.minimum(wi, thld)
    if fix_edge:
        rho = np.sum(pts.astype(np.float32) ** 2, axis=-1)
        good = rho > 0.6 * np.max(rho)
        poly = np.polynomial.Polynomial.fit(rho[good], wi[good], 3)
        wi = wi.copy()
        wi[wi == 0] = poly(rho[wi == 0])
    return wi.astype(np.float32)


def normalize_weight(w):
    s = np.sum(w)
    w = np.where(np.isclose(w, 0.0), 1.0, w).astype(np.float32)
    inv = s / w
    inv = inv / np.sum(inv)
    inv = np.where(np.isclose(inv, 0.0), 1.0, inv)
    return (1 / inv).astype(np.float32)


def voronoi(coord, nshots, fullspoke, fix_edge=False, ratio=0.1):
    c = np.asarray(coord)
    if np.iscomplexobj(c):
        c = np.stack([c.real, c.imag], -1)
    if c.dtype != np.float64:
        c = c.astype(np.float32)
    if nshots > 1:
        kshot = c.reshape(-1, *c.shape[-2:])[0].astype(np.float64)
        ang = np.mod(np.arange(nshots) * (2 * math.pi / nshots * (1.0 if fullspoke else 0.5)), 2 * np.pi)
        cs, sn = np.cos(ang), np.sin(ang)
        x = cs[:, None] * kshot[None, :, 0] - sn[:, None] * kshot[None, :, 1]
        y = sn[:, None] * kshot[None, :, 0] + cs[:, None] * kshot[None, :, 1]
        c = np.stack([x, y], -1).astype(np.float32)
        k0 = int(np.argmin(np.sum(c[0] ** 2, -1) ** 0.5))
    else:
        k0 = int(np.argmin((np.sum(c ** 2, -1) ** 0.5).reshape(-1)))
    ishape = c.shape[:-1]
    pts = c.reshape(-1, 2)
    uniq, inv, counts = np.unique(pts, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    wu = cell_areas(uniq).astype(np.float32)
    wu = _clip_and_fix(wu, uniq, fix_edge, ratio)
    w = (wu / counts.astype(np.float32))[inv]
    w = np.nan_to_num(w, posinf=0.0, neginf=0.0).reshape(ishape)
    if nshots > 1:
        w = w[0]
    w = w.squeeze()
    w = normalize_weight(w)
    return (w * (1 / w.reshape(-1)[k0] / nshots)).astype(np.float32)
