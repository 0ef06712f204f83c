"""Trajectory normalisation.

Intent of src/pyfourier/_subroutines/_traj.py:10-54: map coordinates of any
unit (1/m, rad, [-0.5, 0.5], [-pi, pi)) to grid units [-N/2, N/2).  The
reference divides every point by its *own* radius (defect A.11), which
projects all samples onto one circle; here one global radius (the largest
|k| of the trajectory) is mapped to N/2, so [-pi, pi) input becomes k*N/(2 pi)
and input already in grid units with its largest radius at N/2 is unchanged.
"""

from __future__ import annotations

import torch


def normalize_coordinates(coord: torch.Tensor, shape_xyz, center: bool = True) -> torch.Tensor:
    """Return float32 coordinates in grid units.

    coord: (..., d) with the last axis ordered (x, y, z).
    shape_xyz: matrix size per axis, ordered like the last axis of coord.
    """
    coord = coord.to(torch.float32)
    shape = torch.as_tensor(list(shape_xyz), dtype=torch.float32, device=coord.device)
    cabs = coord.double().pow(2).sum(dim=-1).sqrt().max()
    if float(cabs) == 0.0:
        return coord.clone()
    cabs = cabs.to(torch.float32)
    if bool((coord < 0).any()):
        out = (0.5 * coord / cabs) * shape
        if not center:
            out = out + torch.floor(shape / 2)
        return out
    out = (coord / cabs) * shape
    if center:
        out = out - torch.floor(shape / 2)
    return out
