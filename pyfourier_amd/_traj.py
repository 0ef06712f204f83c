"""Trajectory normalisation.

Intent of src/pyfourier/_subroutines/_traj.py:10-54: accept coordinates in
any of the usual units and return grid units [-N/2, N/2).  The reference
divides every point by its *own* radius (defect A.11), which projects all
samples onto the circle |k| = N/2.  Here the unit is detected once for the
whole trajectory from its largest radius r = max |k|:

* r <= 0.5 (+1e-5)      cycles per field of view [-0.5, 0.5): k * N
* r <= pi  (+1e-5)      radians [-pi, pi):                    k * N / (2 pi)
* otherwise             already grid units:                   unchanged

so radial and spiral trajectories in radians map exactly like the reference
intends (largest radius pi -> N/2) and trajectories already in grid units,
including mixed Cartesian/non-Cartesian ones (stack of stars), pass through
unchanged.
"""

from __future__ import annotations

import math

import torch

_TOL = 1e-5


def coordinate_scale(rmax: float) -> float:
    """Factor (per unit of N) that maps the detected unit to grid units."""
    if rmax <= 0.5 + _TOL:
        return 1.0
    if rmax <= math.pi + _TOL:
        return 1.0 / (2.0 * math.pi)
    return 0.0  # grid units


def unit_factors(coord: torch.Tensor, shape_xyz) -> tuple:
    """d(grid-unit coordinate)/d(input coordinate) per axis (x, y, z order):
    N_a * f for the detected unit, 1 for coordinates already in grid units."""
    rmax = float(coord.detach().double().pow(2).sum(dim=-1).sqrt().max()) if coord.numel() else 0.0
    f = coordinate_scale(rmax)
    return tuple(1.0 if f == 0.0 else float(n) * f for n in shape_xyz)


def normalize_coordinates(coord: torch.Tensor, shape_xyz, center: bool = True) -> torch.Tensor:
    """Return float32 coordinates in grid units.

    coord: (..., d) with the last axis ordered (x, y, z).
    shape_xyz: matrix size per axis, ordered like the last axis of coord.
    center: False shifts the result to [0, N) (reference keyword).
    """
    coord = coord.to(torch.float32)
    shape = torch.as_tensor(list(shape_xyz), dtype=torch.float32, device=coord.device)
    rmax = float(coord.double().pow(2).sum(dim=-1).sqrt().max()) if coord.numel() else 0.0
    f = coordinate_scale(rmax)
    out = coord.clone() if f == 0.0 else coord * (shape * f)
    if not center:
        out = out + torch.floor(shape / 2)
    return out
