"""NUFFT planning: plan_nufft and NUFFTPlan.

Mirrors src/pyfourier/_nufft/_plan.py:20-225.  The Python part normalises the
trajectory, detects Cartesian axes and builds the field-map basis; the
interpolator itself (tap records, tile binning, sort, spread work lists) is
built by the native engine on the GPU (pf_plan_create) the first time the
plan is moved to a device.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._b0 import mri_exp_approx
from ._traj import check_finite, normalize_coordinates, unit_factors


def resolve_device(device) -> torch.device:
    """Map the reference's device argument ("cpu", "cuda:n", int) to a torch
    device.  There is no CPU product path: "cpu" (or -1) raises."""
    if isinstance(device, torch.device):
        dev = device
    elif isinstance(device, str):
        dev = torch.device(device)
    elif isinstance(device, int):
        dev = torch.device("cpu") if device < 0 else torch.device("cuda", device)
    else:
        raise TypeError(f"unsupported device specification {device!r}")
    if dev.type != "cuda":
        raise RuntimeError(
            "pyfourier_amd runs on the GPU only (hand-written HIP kernels, no CPU path): "
            f"pass device='cuda:n' instead of {device!r}"
        )
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


class NativePlan:
    """Owns one pf_plan on one device."""

    def __init__(self, coord_grid: torch.Tensor, nframes: int, npts: int, shape, width, oversamp, device,
                 tile=None, chunk=0, frame_offsets=None):
        """frame_offsets (nframes + 1 ints): a ragged plan whose frame f owns
        samples frame_offsets[f]..frame_offsets[f+1]-1 (npts is then ignored)."""
        self.device = device
        self.ragged = frame_offsets is not None
        d = len(shape)
        c = coord_grid.detach().to(device=device, dtype=torch.float32).reshape(-1, d).contiguous()
        L = _lib.lib()
        sh = (ctypes.c_int * d)(*[int(s) for s in shape])
        wi = (ctypes.c_int * d)(*[int(w) for w in width])
        ov = (ctypes.c_float * d)(*[float(o) for o in oversamp])
        handle = ctypes.c_void_p()
        ti = None if tile is None else (ctypes.c_int * d)(*[int(t) for t in tile])
        with torch.cuda.device(device):
            if self.ragged:
                off = [int(o) for o in frame_offsets]
                if len(off) != nframes + 1 or off[-1] != c.shape[0]:
                    raise ValueError("frame_offsets must have nframes + 1 entries ending at the sample count")
                _lib.check(
                    L.pf_plan_create_ragged(
                        _lib.ptr(c), d, sh, wi, ov, int(nframes), (ctypes.c_longlong * len(off))(*off), ti,
                        int(chunk), _lib.stream_of(device), ctypes.byref(handle),
                    )
                )
            else:
                _lib.check(
                    L.pf_plan_create_tiled(
                        _lib.ptr(c), d, sh, wi, ov, int(nframes), int(npts), ti, int(chunk),
                        _lib.stream_of(device), ctypes.byref(handle),
                    )
                )
        self.handle = handle
        info = _lib.PlanInfo()
        _lib.check(L.pf_plan_get_info(handle, ctypes.byref(info)))
        self.info = info
        self._lib = L
        self._perm = None

    def permutation(self) -> torch.Tensor:
        """Original sample index (frame * npts + j) of every tile-sorted position."""
        if self._perm is None:
            n = int(self.info.npts) if self.ragged else int(self.info.nframes) * int(self.info.npts)
            perm = torch.empty(n, dtype=torch.int32, device=self.device)
            with torch.cuda.device(self.device):
                _lib.check(self._lib.pf_plan_get_permutation(self.handle, _lib.ptr(perm), _lib.stream_of(self.device)))
            self._perm = perm.long()
        return self._perm

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and self._lib is not None:
            try:
                self._lib.pf_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


@dataclass
class Interpolator:
    """Trajectory in grid units plus the native plans built from it.

    Stands in for src/pyfourier/_subroutines/_grid/_plan.py:14-114 (which
    holds per-axis index/value tables); the tables live inside the native
    plan as compact per-sample tap records in tile-sorted order.
    """

    coord: torch.Tensor  # (F, nviews, nsamples, d) grid units, float32
    shape: tuple  # (z, y, x)
    width: tuple
    oversamp: tuple
    beta: tuple
    os_shape: tuple
    scale: float
    ndim: int
    dshape: tuple  # (nviews, nsamples)
    nframes: int
    unit_scale: tuple = ()  # d(grid units)/d(input units) per coord axis (x, y, z)
    tile: tuple = None  # sort/spread tile (z, y, x); None: the engine's default
    chunk: int = 0  # spread-list entries per work item (0: default)
    _native: dict = field(default_factory=dict, repr=False)

    @property
    def npts(self) -> int:
        return int(np.prod(self.dshape))

    def native(self, device: torch.device) -> NativePlan:
        key = (device.type, device.index)
        if key not in self._native:
            self._native[key] = NativePlan(
                self.coord, self.nframes, self.npts, self.shape, self.width, self.oversamp, device, self.tile,
                self.chunk,
            )
        return self._native[key]


@dataclass
class NUFFTPlan:
    """Same fields as the reference NUFFTPlan (src/pyfourier/_nufft/_plan.py:203-215)."""

    ndim: int
    oversamp: tuple
    width: tuple
    beta: tuple
    os_shape: tuple
    shape: tuple
    interpolator: Interpolator
    zmap_t_kernel: object
    zmap_s_kernel: object
    zmap_batch_size: int
    device: object

    def to(self, device):
        """Build (once per device) the native plan on `device` and move the
        field-map basis there (src/pyfourier/_nufft/_plan.py:217-225)."""
        dev = resolve_device(device)
        self.interpolator.native(dev)
        if self.zmap_s_kernel is not None:
            self.zmap_t_kernel = self.zmap_t_kernel.to(dev)
            self.zmap_s_kernel = self.zmap_s_kernel.to(dev)
        self.device = dev
        return self


def _per_axis(v, ndim, dtype):
    if np.isscalar(v):
        return np.asarray([v] * ndim, dtype=dtype)
    return np.asarray(list(v)[-ndim:], dtype=dtype)


def kb_beta(width, oversamp) -> np.ndarray:
    """beta = pi*sqrt((W/s*(s-1/2))^2 - 0.8) in float32 (src/pyfourier/_nufft/_plan.py:128)."""
    w = np.asarray(width, dtype=np.int16)
    s = np.asarray(oversamp, dtype=np.float32)
    inner = (((w / s) * (s - np.float32(0.5))) ** 2 - np.float32(0.8)).astype(np.float32)
    with np.errstate(invalid="ignore"):  # NaN on width-1 (Cartesian) axes, unused there
        return (np.float32(math.pi) * inner ** np.float32(0.5)).astype(np.float32)


def _kb_host(x, beta):
    # host twin of the engine's KB evaluation, used only for the plan summary scale
    x = float(x)
    if abs(x) > 1:
        return 0.0
    u = float(beta) * math.sqrt(1 - x * x)
    t = u / 3.75
    if u < 3.75:
        t2 = t * t
        return 1 + t2 * (3.5156229 + t2 * (3.0899424 + t2 * (1.2067492 + t2 * (0.2659732 + t2 * (0.0360768 + t2 * 0.0045813)))))
    it = 1 / t
    p = 0.39894228 + it * (0.01328592 + it * (0.00225319 + it * (-0.00157565 + it * (0.00916281 + it * (-0.02057706 + it * (0.02635537 + it * (-0.01647633 + it * 0.00392377)))))))
    return p * math.exp(u) / math.sqrt(u)


def _kernel_scale(beta, width) -> float:
    s = 1.0
    for b, w in zip(beta, width):
        w = int(w)
        if w == 1:
            continue
        start = math.ceil(-w / 2)
        s *= sum(_kb_host((start + e) / (w / 2), b) for e in range(w))
    return s


def plan_nufft(
    coord,
    shape,
    width=4,
    oversamp=1.25,
    zmap=None,
    L=6,
    nbins=(40, 40),
    dt=None,
    T=None,
    L_batch_size=None,
    tile=None,
    chunk=0,
    units="auto",
):
    """Precompute a NUFFT plan (same signature as the reference plan_nufft).

    coord: (ncontrasts, nviews, nsamples, ndim) in any unit, last axis (x, y[, z]);
    shape: image matrix (scalar or (z, y, x)); width/oversamp: scalar or per axis (z, y, x).
    tile, chunk (extensions): the sort/spread tile (z, y, x) and the spread
    work-item size, e.g. from pyfourier_amd.autotune_plan; None/0 = defaults.
    units (extension): "auto" (detect), "grid", "radians", "cycles" or "max"
    (any other unit, largest |k| = k-space edge); see normalize_coordinates.
    Non-finite coordinates raise ValueError.
    """
    coord = torch.as_tensor(coord)
    if torch.is_complex(coord):
        raise TypeError("coordinates must be real")
    check_finite(coord)
    ndim = int(coord.shape[-1])
    if ndim not in (1, 2, 3):
        raise ValueError("coordinates must have 1, 2 or 3 spatial dims")
    while coord.dim() < 4:
        coord = coord[None]
    width = _per_axis(width, ndim, np.int16)
    oversamp = _per_axis(oversamp, ndim, np.float32)
    shape = _per_axis(shape, ndim, np.int64)
    unit_scale = unit_factors(coord, shape[::-1].tolist(), units)
    coord = normalize_coordinates(coord, shape[::-1].tolist(), units=units)
    # Cartesian axes get oversamp 1 and width 1 (src/pyfourier/_nufft/_plan.py:139-146)
    is_cart = np.asarray(
        [bool(torch.allclose(coord[..., a], torch.round(coord[..., a]))) for a in range(ndim)][::-1]
    )
    oversamp[is_cart] = 1.0
    width[is_cart] = 1
    beta = kb_beta(width, oversamp)
    os_shape = np.ceil(oversamp * shape.astype(np.float32)).astype(np.int64)
    interp = Interpolator(
        coord=coord.contiguous(),
        shape=tuple(int(s) for s in shape),
        width=tuple(int(w) for w in width),
        oversamp=tuple(float(s) for s in oversamp),
        beta=tuple(float(b) for b in beta),
        os_shape=tuple(int(s) for s in os_shape),
        scale=_kernel_scale(beta[::-1], width[::-1]),
        ndim=ndim,
        dshape=tuple(int(s) for s in coord.shape[1:-1]),
        nframes=int(coord.shape[0]),
        unit_scale=unit_scale,
        tile=None if tile is None else tuple(int(t) for t in tile),
        chunk=int(chunk),
    )
    if zmap is not None:
        if T is None:
            if dt is None:
                raise ValueError("Please provide raster time dt if T is not known")
            T = dt * torch.arange(coord.shape[-2], dtype=torch.float32)
        zt, zs = mri_exp_approx(zmap, T, L, nbins)
        if L_batch_size is None:
            # an x grid with a factor 7 has register FFT passes only, which
            # take at most 16 segments per call (batching segments is exact)
            L_batch_size = min(L, 16) if int(os_shape[-1]) % 7 == 0 else L
    else:
        zt, zs = None, None
    return NUFFTPlan(
        ndim,
        tuple(float(s) for s in oversamp),
        tuple(int(w) for w in width),
        tuple(float(b) for b in beta),
        tuple(int(s) for s in os_shape),
        tuple(int(s) for s in shape),
        interp,
        zt,
        zs,
        L_batch_size,
        None,
    )
