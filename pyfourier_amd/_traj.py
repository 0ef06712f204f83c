"""Trajectory normalisation.

Intent of src/pyfourier/_subroutines/_traj.py:10-54: accept coordinates in
any of the usual units (1/m, rad/m, cycles/FOV [-0.5, 0.5), radians
[-pi, pi)) and return grid units [-N/2, N/2).  The reference divides every
point by its *own* radius (defect A.11), which projects all samples onto the
circle |k| = N/2.  Here one factor is applied to the whole trajectory.

``units`` names the input unit (m = the largest |k_a| over all samples and
axes, so a square k-space is measured by its edge, not its corners):

* ``"cycles"``    cycles per field of view [-0.5, 0.5):  k * N
* ``"radians"``   radians [-pi, pi):                     k * N / (2 pi)
* ``"grid"``      already grid units [-N/2, N/2):        unchanged
* ``"max"``       any unit (1/m, rad/m, ...): m is taken as the k-space
                  edge:                                  k * N / (2 m)
* ``"auto"``      (default) detected from m and the matrix size N:
                  m <= 0.5 -> cycles, m <= pi -> radians,
                  |k_a| <= N_a/2 on every axis -> grid, otherwise -> max
                  (physical units exceed N/2 whenever the resolution is finer
                  than one unit of length per pixel).

The detection cannot tell a grid-unit trajectory with m <= pi (a
calibration region, a matrix of a few pixels) from one in radians; such a
trajectory is taken as radians, as the reference would, and callers with
grid-unit coordinates say ``units="grid"``.  Non-finite coordinates are
rejected with ValueError before any planning.
"""

from __future__ import annotations

import math

import torch

_TOL = 1e-5
UNITS = ("auto", "grid", "radians", "cycles", "max")


def _amax(coord: torch.Tensor) -> list:
    """Largest |k_a| per axis (the last axis of coord)."""
    if not coord.numel():
        return [0.0] * int(coord.shape[-1])
    return [float(v) for v in coord.detach().abs().reshape(-1, coord.shape[-1]).max(dim=0).values.double()]


def check_finite(coord: torch.Tensor) -> None:
    """Raise ValueError if any coordinate is NaN or infinite."""
    if coord.numel() and not bool(torch.isfinite(coord.detach()).all()):
        raise ValueError("coordinates must be finite (found NaN or Inf)")


def resolve_units(units: str, amax, shape_xyz) -> str:
    """The unit `units` stands for, given the largest |k_a| per axis and the
    matrix (both ordered like the coordinate axis)."""
    if units not in UNITS:
        raise ValueError(f"units must be one of {UNITS}, got {units!r}")
    if units != "auto":
        return units
    m = max(amax) if len(amax) else 0.0
    if m <= 0.5 + _TOL:
        return "cycles"
    if m <= math.pi + _TOL:
        return "radians"
    if all(a <= 0.5 * float(n) * (1 + 1e-6) + _TOL for a, n in zip(amax, shape_xyz)):
        return "grid"
    return "max"


def _per_n(units: str, amax, shape_xyz) -> float:
    """Factor per unit of N that maps `units` to grid units (0: unchanged)."""
    u = resolve_units(units, amax, shape_xyz)
    m = max(amax) if len(amax) else 0.0
    if u == "cycles":
        return 1.0
    if u == "radians":
        return 1.0 / (2.0 * math.pi)
    if u == "max" and m > 0.0:
        return 0.5 / m
    return 0.0


def unit_factors(coord: torch.Tensor, shape_xyz, units: str = "auto") -> tuple:
    """d(grid-unit coordinate)/d(input coordinate) per axis (x, y, z order).

    For ``"max"`` the edge m is held fixed (it selects the scale, as the
    detected unit does), so the factor is N_a / (2 m)."""
    f = _per_n(units, _amax(coord), shape_xyz)
    return tuple(1.0 if f == 0.0 else float(n) * f for n in shape_xyz)


def normalize_coordinates(coord: torch.Tensor, shape_xyz, center: bool = True, units: str = "auto") -> torch.Tensor:
    """Return float32 coordinates in grid units.

    coord: (..., d) with the last axis ordered (x, y, z).
    shape_xyz: matrix size per axis, ordered like the last axis of coord.
    center: False shifts the result to [0, N) (reference keyword).
    units: input unit (module docstring); "auto" detects it.
    """
    coord = torch.as_tensor(coord)
    if torch.is_complex(coord):
        raise TypeError("coordinates must be real")
    check_finite(coord)
    coord = coord.to(torch.float32)
    shape = torch.as_tensor(list(shape_xyz), dtype=torch.float32, device=coord.device)
    f = _per_n(units, _amax(coord), shape_xyz)
    out = coord.clone() if f == 0.0 else coord * (shape * f)
    if not center:
        out = out + torch.floor(shape / 2)
    return out
