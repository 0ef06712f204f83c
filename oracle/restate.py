"""CPU restatement of pyfourier's NUFFT path (numpy).

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and by nothing else.  The product (``pyfourier_amd``) never
imports, calls or links it.

It restates the *intended* algorithm of the reference (paths relative to the
reference repository root):

* plan parameters      src/pyfourier/_nufft/_plan.py:117-155
* KB kernel            src/pyfourier/_subroutines/_grid/_plan.py:157-189
* tap tables           src/pyfourier/_subroutines/_grid/_plan.py:14-91,192-210
* kernel scale         src/pyfourier/_subroutines/_grid/_plan.py:118-141
* coordinate scaling   src/pyfourier/_subroutines/_traj.py:57-66
* normalisation        src/pyfourier/_subroutines/_traj.py:10-54 (intent: one
                       global radius, not each point's own radius)
* apodisation          src/pyfourier/_subroutines/_apod.py:16-46
* zero-pad / crop      src/pyfourier/_subroutines/_resize.py:16-42
* centered FFT         src/pyfourier/_subroutines/_fft.py:72-79,163-170
* gather               src/pyfourier/_subroutines/_grid/_cart2noncart/_{1,2,3}D.py
* spread               src/pyfourier/_subroutines/_grid/_noncart2cart/_{1,2,3}D.py
* subspace gather      src/pyfourier/_subroutines/_grid/_cart2noncart/_2D_subspace.py:13-54
* subspace spread      src/pyfourier/_subroutines/_grid/_noncart2cart/_2D_subspace.py:13-53
* operator pipeline    src/pyfourier/_nufft/_nufft.py:11-169,218-287
* B0 segmentation      src/pyfourier/_subroutines/_b0.py:10-135
* coordinate gradient  src/pyfourier/_nufft/_nufft.py:175-209 (intent)
* Pipe-Menon DCF       src/pyfourier/_dcomp/pipemenon.py:83-115
* Toeplitz Gram        src/pyfourier/_sparsegram/_plan.py:108-249,
                       src/pyfourier/_sparsegram/_sparsegram.py:11-81
* Cartesian FFT        src/pyfourier/_fft/_fft.py:10-151 (masked FFT with
                       subspace basis and B0 segments along the readout)
* dense Gram           src/pyfourier/_gram/_plan.py:124-160 (kernel),
                       src/pyfourier/_gram/_gram.py:11-133 (apply), intent
* sparse FFT           src/pyfourier/_sparsefft/_sparsefft.py:11-144 with the
                       pick / zero-fill of src/pyfourier/_subroutines/_mask/
* GROG                 src/pyfourier/_preinterp/_grog.py:22-425 (intent: the
                       row-vector operator product, one shift per sample)

Arrays follow the reference conventions: coordinates ``(F, M, d)`` with the
last axis ordered (x, y, z); grid/width/oversampling ordered (z, y, x).

How each function is pinned (independently of this file) is listed in
``tests/test_oracle.py``; functions with no independent pin are marked
"parity unpinned" below.
"""

from __future__ import annotations

import math

import numpy as np

# --------------------------------------------------------------------------
# Kaiser-Bessel kernel: I0(beta*sqrt(1-x^2)) with the Abramowitz & Stegun
# 9.8.1 (|u| < 3.75) and 9.8.2 (u >= 3.75) polynomial approximations, zero for
# |x| > 1.  Follows src/pyfourier/_subroutines/_grid/_plan.py:157-189.
# --------------------------------------------------------------------------
_AS_SMALL = (1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.0360768, 0.0045813)
_AS_LARGE = (
    0.39894228,
    0.01328592,
    0.00225319,
    -0.00157565,
    0.00916281,
    -0.02057706,
    0.02635537,
    -0.01647633,
    0.00392377,
)


def kb(x, beta):
    """KB kernel value at normalised offset ``x`` (float64)."""
    x = np.asarray(x, dtype=np.float64)
    beta = float(beta)
    inside = np.abs(x) <= 1.0
    u = beta * np.sqrt(np.where(inside, 1.0 - x * x, 0.0))
    t = u / 3.75
    t2 = t * t
    small = 0.0
    for c in reversed(_AS_SMALL):
        small = small * t2 + c
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        it = np.where(u >= 3.75, 1.0 / np.where(t > 0, t, 1.0), 0.0)
        large = 0.0
        for c in reversed(_AS_LARGE):
            large = large * it + c
        large = np.where(u >= 3.75, u ** -0.5 * np.exp(u) * large, 0.0)
    val = np.where(u < 3.75, small, large)
    return np.where(inside, val, 0.0)


def kb_deriv(x, beta):
    """d kb(x, beta) / dx (float64), 0 for |x| > 1.

    Differentiates the same two polynomial branches as ``kb``: with
    u = beta sqrt(1 - x^2), du^2/dx = -2 beta^2 x and du/dx = -beta^2 x / u.
    Used by the coordinate gradient (the reference's _nufft_backward,
    src/pyfourier/_nufft/_nufft.py:175-209, states the intent only).
    """
    x = np.asarray(x, dtype=np.float64)
    beta = float(beta)
    inside = np.abs(x) <= 1.0
    u2 = beta * beta * np.where(inside, 1.0 - x * x, 0.0)
    u = np.sqrt(u2)
    t2 = u2 / (3.75 * 3.75)
    # small branch: P(t2) = sum_k a_k t2^k -> dP/dx = P'(t2) * dt2/dx
    dp = 0.0
    for k in range(len(_AS_SMALL) - 1, 0, -1):
        dp = dp * t2 + k * _AS_SMALL[k]
    small = dp * (-2.0 * beta * beta * x / (3.75 * 3.75))
    # large branch: Q(u) = u^-1/2 e^u q(3.75/u)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        us = np.where(u >= 3.75, u, 3.75)
        s = 3.75 / us
        q = 0.0
        dq = 0.0
        for k in range(len(_AS_LARGE) - 1, -1, -1):
            q = q * s + _AS_LARGE[k]
            if k > 0:
                dq = dq * s + k * _AS_LARGE[k]
        dQ = us ** -0.5 * np.exp(us) * (q * (1.0 - 0.5 / us) - dq * 3.75 / (us * us))
        large = dQ * (-beta * beta * x / us)
    val = np.where(u < 3.75, small, large)
    return np.where(inside, val, 0.0)


# --------------------------------------------------------------------------
# Plan parameters: src/pyfourier/_nufft/_plan.py:117-155
# --------------------------------------------------------------------------
def kb_beta(width, oversamp):
    """beta = pi*sqrt((W/s*(s-1/2))^2 - 0.8), evaluated in float32 like the
    reference (width int16, oversamp float32): _nufft/_plan.py:128."""
    w = np.asarray(width, dtype=np.int16)
    s = np.asarray(oversamp, dtype=np.float32)
    inner = ((w / s) * (s - np.float32(0.5))) ** 2 - np.float32(0.8)
    with np.errstate(invalid="ignore"):  # NaN on width-1 axes, never used
        return (np.float32(math.pi) * inner.astype(np.float32) ** np.float32(0.5)).astype(np.float32)


def os_shape(shape, oversamp):
    """ceil(sigma*N) per axis: _subroutines/_utils/_common.py:16-18 (intent)."""
    s = np.asarray(oversamp, dtype=np.float32)
    n = np.asarray(shape, dtype=np.int64)
    return [int(v) for v in np.ceil(s * n.astype(np.float32))]


def kernel_scale(beta, width):
    """Product over axes of sum_e KB((ceil(-W/2)+e)/(W/2)); 1 on W=1 axes.
    _subroutines/_grid/_plan.py:118-141."""
    total = 1.0
    for b, w in zip(beta, width):
        w = int(w)
        if w == 1:
            continue
        start = math.ceil(-w / 2)
        total *= float(np.sum([kb((start + e) / (w / 2), b) for e in range(w)]))
    return total


def normalize_coordinates(coord, shape_xyz, units="auto"):
    """Map coordinates to grid units [-N/2, N/2).

    Intent of _subroutines/_traj.py:10-54 ("input can have any normalization,
    e.g. 1/m, rad/m, -0.5:0.5, -pi:pi") with defect A.11 fixed (the reference
    divides each point by its own radius, :37-38,47-48).  One factor for the
    whole trajectory, from the largest |k_a| over samples and axes, m, and the
    matrix N: m <= 0.5 -> cycles/FOV (times N), m <= pi -> radians (times
    N/(2 pi)), |k_a| <= N_a/2 on every axis -> grid units (unchanged),
    otherwise any other unit with m as the k-space edge (times N/(2 m)).
    `units` forces one of "cycles", "radians", "grid", "max".
    """
    coord = np.asarray(coord, dtype=np.float32)
    if not np.isfinite(coord).all():
        raise ValueError("coordinates must be finite")
    shape_xyz = np.asarray(shape_xyz, dtype=np.float32)
    d = coord.shape[-1]
    amax = np.abs(coord.reshape(-1, d)).max(axis=0).astype(np.float64) if coord.size else np.zeros(d)
    m = float(amax.max()) if d else 0.0
    if units == "auto":
        if m <= 0.5 + 1e-5:
            units = "cycles"
        elif m <= math.pi + 1e-5:
            units = "radians"
        elif np.all(amax <= 0.5 * shape_xyz.astype(np.float64) * (1 + 1e-6) + 1e-5):
            units = "grid"
        else:
            units = "max"
    if units == "cycles":
        f = 1.0
    elif units == "radians":
        f = 1.0 / (2.0 * math.pi)
    elif units == "max" and m > 0.0:
        f = 0.5 / m
    elif units in ("grid", "max"):
        return coord.copy()
    else:
        raise ValueError(f"unknown units {units!r}")
    return coord * (shape_xyz * np.float32(f))


def scale_coord(coord, shape_xyz, oversamp_xyz, dtype=np.float32):
    """c_os = c*ceil(sN)/N + ceil(sN)//2 in float32: _subroutines/_traj.py:57-66.
    dtype float64 gives the exact map (finite-difference checks only)."""
    out = np.array(coord, dtype=dtype, copy=True)
    for a in range(out.shape[-1]):
        n = np.float32(shape_xyz[a])
        g = np.ceil(np.float32(oversamp_xyz[a]) * n)
        out[..., a] = out[..., a] * dtype(g / n) + dtype(g // 2)
    return out


def interp_tables(coord_os, width_xyz, beta_xyz, grid_xyz, dtype=np.float32):
    """Per-axis tap index/value tables, ordered (x, y, z).

    index[a] (P, W_a) int32 and value[a] (P, W_a) float32 for P points, from
    _subroutines/_grid/_plan.py:192-210: x0 = ceil(c - W/2), taps x0+i,
    value KB((x0+i-c)/(W/2)) computed in float64 and stored as float32,
    index (x0+i) mod G.  W=1 axes use the integer coordinate with value 1
    (_grid/_plan.py:68-72, intent of defect A.9).
    """
    c = np.asarray(coord_os, dtype=dtype).reshape(-1, coord_os.shape[-1])
    index, value = [], []
    for a in range(c.shape[-1]):
        w = int(width_xyz[a])
        g = int(grid_xyz[a])
        ca = c[:, a].astype(np.float64)
        if w == 1:
            index.append(np.mod(np.rint(ca).astype(np.int64), g).astype(np.int32)[:, None])
            value.append(np.ones((c.shape[0], 1), dtype=dtype))
            continue
        x0 = np.ceil(ca - w / 2)
        taps = x0[:, None] + np.arange(w)[None, :]
        value.append(kb((taps - ca[:, None]) / (w / 2), beta_xyz[a]).astype(dtype))
        index.append(np.mod(taps, g).astype(np.int32))
    return index, value


def apod_vector(n, oversamp, width, beta):
    """KB deapodisation along one axis of length n: _subroutines/_apod.py:27-41.
    The grid length is ceil(sigma * n) in float32 (the reference's float32
    oversampling factor times the axis length, _apod.py:27), as os_shape."""
    if int(width) == 1:
        return np.ones(n, dtype=np.float64)
    os_n = int(np.ceil(np.float32(oversamp) * np.float32(n)))
    idx = np.arange(n, dtype=np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        a = np.sqrt(float(beta) ** 2 - (math.pi * int(width) * (idx - n // 2) / os_n) ** 2)
        a = a / np.sinh(a)
        a = a / a[n // 2]
    return np.nan_to_num(a, nan=1.0)


def resize(x, oshape):
    """Centered zero-pad or crop of the trailing axes: _subroutines/_resize.py:16-42."""
    x = np.asarray(x)
    d = len(oshape)
    ishape = x.shape[-d:]
    lead = x.shape[:-d]
    out = np.zeros(lead + tuple(oshape), dtype=x.dtype)
    isl, osl = [], []
    for i, o in zip(ishape, oshape):
        si = max(i // 2 - o // 2, 0)
        so = max(o // 2 - i // 2, 0)
        cpy = min(i - si, o - so)
        isl.append(slice(si, si + cpy))
        osl.append(slice(so, so + cpy))
    out[(Ellipsis,) + tuple(osl)] = x[(Ellipsis,) + tuple(isl)]
    return out


def fftc(x, ndim, norm=None, inverse=False, centered=True):
    """fftshift(fftn(ifftshift(x))) over the last ndim axes (torch branch of
    _subroutines/_fft.py:72-79 / 163-170); norm None or 'ortho'."""
    axes = tuple(range(-ndim, 0))
    f = np.fft.ifftn if inverse else np.fft.fftn
    if centered:
        return np.fft.fftshift(f(np.fft.ifftshift(x, axes=axes), axes=axes, norm=norm), axes=axes)
    return f(x, axes=axes, norm=norm)


# --------------------------------------------------------------------------
# Plan
# --------------------------------------------------------------------------
class Plan:
    """Restatement of plan_nufft (src/pyfourier/_nufft/_plan.py:20-199).

    ``coord`` (F, ..., d) in any unit; ``shape`` the image matrix (z, y, x).
    """

    def __init__(self, coord, shape, width=4, oversamp=1.25, normalize=True, precise=False):
        # precise: coordinates in grid units kept in float64 end to end (for
        # finite-difference checks of derivatives; not the reference's rounding)
        dt = np.float64 if precise else np.float32
        coord = np.asarray(coord, dtype=dt)
        d = coord.shape[-1]
        while coord.ndim < 4:
            coord = coord[None]
        self.ndim = d
        self.nframes = coord.shape[0]
        self.dshape = coord.shape[1:-1]  # (nviews, nsamples)
        self.npts = int(np.prod(self.dshape))
        shape = [shape] * d if np.isscalar(shape) else list(shape)[-d:]
        width = [width] * d if np.isscalar(width) else list(width)
        oversamp = [oversamp] * d if np.isscalar(oversamp) else list(oversamp)
        self.shape = [int(s) for s in shape]  # (z, y, x)
        shape_xyz = self.shape[::-1]
        if normalize and not precise:
            coord = normalize_coordinates(coord, shape_xyz)
        self.coord = coord  # grid units, (F, nv, ns, d)
        # Cartesian axes (all integer) get sigma=1, W=1 (_nufft/_plan.py:139-146)
        is_cart = [np.allclose(coord[..., a], np.round(coord[..., a])) for a in range(d)][::-1]
        width = np.asarray(width, dtype=np.int16)
        oversamp = np.asarray(oversamp, dtype=np.float32)
        width[np.asarray(is_cart)] = 1
        oversamp[np.asarray(is_cart)] = 1.0
        self.width = [int(w) for w in width]  # (z, y, x)
        self.oversamp = [float(s) for s in oversamp]
        self.beta = [float(b) for b in kb_beta(width, oversamp)]
        self.os_shape = os_shape(self.shape, oversamp)  # (z, y, x)
        self.coord_os = scale_coord(coord, shape_xyz, list(oversamp[::-1]), dt)
        self.index, self.value = interp_tables(
            self.coord_os.reshape(-1, d),
            self.width[::-1],
            self.beta[::-1],
            self.os_shape[::-1],
            dt,
        )
        self.scale = kernel_scale(self.beta[::-1], self.width[::-1])

    # separable tap weights for all points: list over axes (x, y, z)
    def taps(self):
        return self.index, self.value


# --------------------------------------------------------------------------
# Gather / spread kernels (per frame), all batch items at once.
# _cart2noncart/_3D.py:13-51 and _noncart2cart/_3D.py:13-51 (and 1D/2D twins)
# --------------------------------------------------------------------------
def deriv_tables(plan):
    """Per-axis d(tap weight)/d(c_os) tables, ordered (x, y, z): the analytic
    derivative of KB((x0+i-c)/(W/2)) with the taps x0 held fixed, 0 on W=1
    axes (their weight does not depend on c)."""
    c = plan.coord_os.reshape(-1, plan.ndim).astype(np.float64)
    out = []
    for a, (w, b) in enumerate(zip(plan.width[::-1], plan.beta[::-1])):
        if w == 1:
            out.append(np.zeros((c.shape[0], 1)))
            continue
        x0 = np.ceil(c[:, a] - w / 2)
        taps = x0[:, None] + np.arange(w)[None, :]
        out.append(-kb_deriv((taps - c[:, a, None]) / (w / 2), b) / (w / 2))
    return out


def _tap_iter(plan, frame, daxis=None):
    """Yield (flat grid index array, weight array) for every tap combination
    of the points of one frame, in (z, y, x) loop order.  With daxis (an
    index into (x, y, z)) the weights are differentiated along that axis."""
    d = plan.ndim
    sl = slice(frame * plan.npts, (frame + 1) * plan.npts)
    idx = [i[sl] for i in plan.index][::-1]  # (z, y, x)
    val = [v[sl].astype(np.float64) for v in plan.value]
    if daxis is not None:
        val[daxis] = deriv_tables(plan)[daxis][sl]
    val = val[::-1]
    gshape = plan.os_shape
    widths = [i.shape[1] for i in idx]
    for taps in np.ndindex(*widths):
        flat = np.zeros(plan.npts, dtype=np.int64)
        w = np.ones(plan.npts, dtype=np.float64)
        for a in range(d):
            flat = flat * gshape[a] + idx[a][:, taps[a]]
            w = w * val[a][:, taps[a]]
        yield flat, w


def gather(grid, plan, daxis=None):
    """grid (B, F, *G) -> samples (B, F, M); /scale applied.  With daxis the
    derivative of the samples with respect to c_os[daxis] (x, y, z index)."""
    grid = np.asarray(grid)
    b = grid.shape[0]
    out = np.zeros((b, plan.nframes, plan.npts), dtype=np.complex128)
    for f in range(plan.nframes):
        g = grid[:, f].reshape(b, -1)
        for flat, w in _tap_iter(plan, f, daxis):
            out[:, f] += w[None] * g[:, flat]
    return out / plan.scale


def spread(ksp, plan):
    """samples (B, F, M) -> grid (B, F, *G); /scale applied."""
    ksp = np.asarray(ksp)
    b = ksp.shape[0]
    gsize = int(np.prod(plan.os_shape))
    out = np.zeros((b, plan.nframes, gsize), dtype=np.complex128)
    for f in range(plan.nframes):
        for flat, w in _tap_iter(plan, f):
            for bi in range(b):
                out[bi, f] += np.bincount(flat, weights=(w * ksp[bi, f]).real, minlength=gsize)
                out[bi, f] += 1j * np.bincount(flat, weights=(w * ksp[bi, f]).imag, minlength=gsize)
    return out.reshape((b, plan.nframes) + tuple(plan.os_shape)) / plan.scale


def apodize(x, plan):
    """Multiply the trailing ndim axes by the separable apodisation (pure)."""
    out = np.array(x, dtype=np.complex128, copy=True)
    d = plan.ndim
    for n in range(1, d + 1):
        ax = -n
        a = apod_vector(out.shape[ax], plan.oversamp[ax], plan.width[ax], plan.beta[ax])
        out = out * a.reshape([-1] + [1] * (n - 1))
    return out


# --------------------------------------------------------------------------
# Operators: src/pyfourier/_nufft/_nufft.py:218-287 (no B0) with the B0 loop
# of :11-169 (defect A.13 fixed: each segment added once).
# Batch layout here: x (B, F|K, *N), y (B, F, M).
# --------------------------------------------------------------------------
def _basis_c(basis):
    return None if basis is None else np.asarray(basis, dtype=np.complex128)


def _to_grid(x, plan, norm):
    """x (B, P, *N) -> oversampled centred k-space grid (B, P, *G)."""
    y = apodize(x, plan)
    y = resize(y, plan.os_shape)
    return fftc(y, plan.ndim, norm=norm)


def _from_grid(g, plan, norm):
    """grid (B, P, *G) -> image (B, P, *N)."""
    x = fftc(g, plan.ndim, norm=norm, inverse=True)
    x = resize(x, plan.shape)
    return apodize(x, plan)


def nufft(x, plan, basis=None, weight=None, norm=None, b0=None, daxis=None):
    """Forward NUFFT.  x (B, F|K, *N) -> y (B, F, M).

    Without basis, frame f of x is sampled with frame f's taps.  With
    basis (F, K): y[f] = sum_k B[f,k] A_f X_k (B.8).  With
    b0 = (Bt (L, nsamples), Ct (L, *N)): y = sum_l Bt[l,t] A(Ct_l x) (B.9).
    With daxis (x, y, z index) returns dy/dc_os[daxis] instead; norm may
    also be numpy's 'forward' (1/G on the FFT, the scaling of the adjoint).
    """
    x = np.asarray(x, dtype=np.complex128)
    if x.ndim != plan.ndim + 2:
        raise ValueError(f"x must be (B, F|K, *N), got shape {x.shape}")
    B = x.shape[0]
    F = plan.nframes
    basis = _basis_c(basis)
    segs = [(None, None)] if b0 is None else list(zip(b0[0], b0[1]))
    y = np.zeros((B, F, plan.npts), dtype=np.complex128)
    for bt, ct in segs:
        xs = x if ct is None else x * ct
        if basis is None:
            ys = gather(_to_grid(xs, plan, norm), plan, daxis)
        else:
            ys = np.zeros((B, F, plan.npts), dtype=np.complex128)
            for k in range(basis.shape[1]):
                gk = _to_grid(xs[:, k : k + 1], plan, norm)
                gk = np.broadcast_to(gk, (B, F) + gk.shape[2:])
                ys += basis[None, :, k, None] * gather(gk, plan, daxis)
        if bt is not None:
            t = np.asarray(bt).reshape(-1)
            ys = ys * np.tile(t, plan.npts // t.size)[None, None]
        y += ys
    if weight is not None:
        y = y * np.asarray(weight).reshape(1, F, -1)
    return y


def nufft_adj(y, plan, basis=None, weight=None, norm=None, b0=None):
    """Adjoint NUFFT.  y (B, F, M) -> x (B, F|K, *N)."""
    y = np.asarray(y, dtype=np.complex128)
    if y.ndim != 3:
        raise ValueError(f"y must be (B, F, M), got shape {y.shape}")
    F = plan.nframes
    if weight is not None:
        y = y * np.asarray(weight).reshape(1, F, -1)
    basis = _basis_c(basis)
    segs = [(None, None)] if b0 is None else list(zip(b0[0], b0[1]))
    x = None
    for bt, ct in segs:
        ys = y
        if bt is not None:
            t = np.asarray(bt).reshape(-1)
            ys = ys * np.conj(np.tile(t, plan.npts // t.size))[None, None]
        if basis is None:
            xs = _from_grid(spread(ys, plan), plan, norm)
        else:
            xs = np.concatenate(
                [
                    _from_grid(
                        spread(np.conj(basis[None, :, k, None]) * ys, plan).sum(axis=1, keepdims=True),
                        plan,
                        norm,
                    )
                    for k in range(basis.shape[1])
                ],
                axis=1,
            )
        if ct is not None:
            xs = xs * np.conj(ct)
        x = xs if x is None else x + xs
    return x


def coord_grad(x, plan, h, basis=None, weight=None, norm=None, b0=None, adjoint=False):
    """Trajectory gradient Re(sum_B conj(h) dy/dk) per sample, (F*M, d) in
    grid units ordered (x, y, z).

    Forward operator y = nufft(x): h is the upstream k-space gradient.
    Adjoint operator x = nufft_adj(y): pass x = the upstream image gradient,
    h = y and adjoint=True (the derivative then carries the adjoint's FFT
    scaling).  Intent of src/pyfourier/_nufft/_nufft.py:175-209 (whose
    ramp-based formula drops the -2 pi i / N factor, defect A.14).
    """
    fnorm = ("forward" if norm is None else norm) if adjoint else norm
    h = np.asarray(h, dtype=np.complex128)
    if weight is not None:
        w = np.asarray(weight).reshape(1, plan.nframes, -1)
        h = h * (w if adjoint else np.conj(w))
    out = np.zeros((plan.nframes * plan.npts, plan.ndim))
    shape_xyz = plan.shape[::-1]
    os_xyz = plan.os_shape[::-1]
    for a in range(plan.ndim):
        dy = nufft(x, plan, basis=basis, norm=fnorm, b0=b0, daxis=a)
        g = np.real(np.conj(h) * dy).sum(axis=0).reshape(-1)
        out[:, a] = g * os_xyz[a] / shape_xyz[a]
    return out


def pipe_menon_dcf(coord, shape, niter=30, width=6, oversamp=1.5):
    """Pipe-Menon density compensation, src/pyfourier/_dcomp/pipemenon.py:83-115:
    w = 1, then niter times w <- w / |G G^H w| with the ortho NUFFT, each
    operator's output cast back to real (its real part), as the reference's
    float dtype round trip does.  Returns w of shape coord.shape[:-1]."""
    P = Plan(coord, shape, width=width, oversamp=oversamp)
    w = np.ones((1, P.nframes, P.npts))
    for _ in range(niter):
        x = nufft_adj(w.astype(np.complex128), P, norm="ortho").real
        y = nufft(x.astype(np.complex128), P, norm="ortho").real
        w = w / np.abs(y)
    return w.reshape(np.asarray(coord).shape[:-1])


# --------------------------------------------------------------------------
# B0 time segmentation: intent of src/pyfourier/_subroutines/_b0.py:10-135
# (MIRT mri_exp_approx 'hist,time,unif'), defect A.15 fixed: bin centres are
# (e[i]+e[i+1])/2 and the temporal basis fits exp(-z t) (same sign as Ct).
# parity unpinned by the reference (it cannot run); pinned instead by the
# approximation error to the exact exp(-z t) (tests/test_oracle.py).
# --------------------------------------------------------------------------
def mri_exp_approx(zmap, t, L=6, nbins=40):
    zmap = np.asarray(zmap)
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    if np.isrealobj(zmap):
        z = 2j * math.pi * zmap.astype(np.float64)
        hk, e = np.histogram(z.imag.ravel(), bins=nbins if np.isscalar(nbins) else nbins[0])
        zk = 1j * (e[:-1] + e[1:]) / 2
    else:
        z = 2 * math.pi * zmap.astype(np.complex128)
        nb = (nbins, nbins) if np.isscalar(nbins) else tuple(nbins)
        hk, ei, er = np.histogram2d(z.imag.ravel(), z.real.ravel(), bins=nb)
        zk = (1j * ((ei[:-1] + ei[1:]) / 2)[:, None] + ((er[:-1] + er[1:]) / 2)[None, :])
    hk = hk.ravel().astype(np.float64)
    zk = zk.ravel()
    tl = np.linspace(0, L, L) / L * t[-1]
    ch = np.exp(-tl[:, None] * zk[None, :])  # (L, nk)
    w = np.sqrt(hk)
    p = np.linalg.pinv(w[:, None] * ch.T) * w[None, :]  # (L, nk)
    bt = p @ np.exp(-zk[:, None] * t[None, :])  # (L, nt)
    ct = np.exp(-tl.reshape((-1,) + (1,) * z.ndim) * z[None])
    return bt, ct


# --------------------------------------------------------------------------
# Toeplitz Gram: B.10.  PSF(dr) = sum_j w_j exp(+2 pi i k_j . dr / N) on the
# 2N grid, T = FFT_u(wrap(PSF)) / Ghat, Gx = crop(IFFT_u(T * FFT_u(pad(x)))).
# ---------------------------------------------------------------------------
def toeplitz_kernel_exact(coord_grid, shape, weight=None, ghat=1.0):
    """FP64 exact Toeplitz kernel from the NUDFT PSF (small problems only)."""
    from .nudft import psf_exact

    psf = psf_exact(coord_grid, shape, weight)  # (2N...) centred at index N
    d = len(shape)
    axes = tuple(range(-d, 0))
    return np.fft.fftn(np.fft.ifftshift(psf, axes=axes), axes=axes) / ghat


def toeplitz_apply(x, T, shape):
    """Gx = crop_N(IFFT_u(T * FFT_u(pad_2N(x)))): _sparsegram/_sparsegram.py:59-81."""
    d = len(shape)
    big = [2 * s for s in shape]
    xp = resize(np.asarray(x, dtype=np.complex128), big)
    yp = fftc(T * fftc(xp, d, centered=False), d, inverse=True, centered=False)
    return resize(yp, list(shape))


def toeplitz_kernel_block_exact(coord_grid, shape, coefs, weight=None, ghat=1.0):
    """FP64 exact per-voxel R x R Toeplitz kernel of the subspace / field-map
    Gram, (R, R, 2N...):

        T[r, r'] = FFT_u(wrap(PSF_rr')) / ghat,
        PSF_rr'(dr) = sum_j w_j conj(c_r(j)) c_r'(j) exp(+2 pi i k_j . dr / N)

    coefs (M, R): each sample's coefficient c_r(j).  Subspace: c_k(j) =
    basis[frame(j), k] (the basis-weighted delta of
    src/pyfourier/_sparsegram/_plan.py:148-156, adjoint-transformed with the
    basis so T[k, k'] = sum_f conj(B[f, k]) B[f, k'] PSF_f); field map: c_l(j)
    = Bt[l, t(j)] (:175-183); both: c_(l, k) = Bt[l, t] basis[f, k], r = l*K + k.
    The kernel layout of :228-239 (one R x R matrix per voxel) with the exact
    A^H W A scale (defect A.16: the reference divides by mean|T|).
    """
    from .nudft import nudft_adj

    coord = np.asarray(coord_grid, dtype=np.float64).reshape(-1, len(shape))
    c = np.asarray(coefs, dtype=np.complex128).reshape(coord.shape[0], -1)
    R_ = c.shape[1]
    w = np.ones(coord.shape[0]) if weight is None else np.asarray(weight, dtype=np.float64).reshape(-1)
    y = w[None, None, :] * np.conj(c.T)[:, None, :] * c.T[None, :, :]  # (R, R, M)
    big = [2 * n for n in shape]
    psf = nudft_adj(y.reshape(R_ * R_, -1), 2.0 * coord, big).reshape((R_, R_) + tuple(big))
    d = len(shape)
    axes = tuple(range(-d, 0))
    return np.fft.fftn(np.fft.ifftshift(psf, axes=axes), axes=axes) / ghat


def toeplitz_apply_block(x, T, shape, b0_spatial=None):
    """Subspace / field-map Toeplitz Gram of one batch item.

    x (K, *N) -> (K, *N).  T (R, R, 2N...) from toeplitz_kernel_block_exact,
    R = L*K; b0_spatial (L, *N) spatial maps Ct_l or None (L = 1).
      u_(l,k) = FFT_u(pad_2N(Ct_l * x_k))                 (_sparsegram.py:26-53, 59-73)
      v_r     = sum_r' T[r, r'] u_r'   per voxel          (_bmatvec/_bdot.py:130-157)
      out_k   = sum_l conj(Ct_l) * crop_N(IFFT_u(v_(l,k))) (_sparsegram.py:44-51, 75-81)
    The per-voxel product is the matrix-vector product the Gram needs; the
    reference's _dot_nb (_bdot.py:119-127) indexes the vector by the output
    index j where the product needs the summation index i.
    """
    x = np.asarray(x, dtype=np.complex128)
    d = len(shape)
    K = x.shape[0]
    big = [2 * n for n in shape]
    maps = np.ones((1,) + tuple(shape)) if b0_spatial is None else np.asarray(b0_spatial, dtype=np.complex128)
    L = maps.shape[0]
    u = np.stack([fftc(resize(maps[l] * x[k], big), d, centered=False) for l in range(L) for k in range(K)])
    v = np.einsum("rs...,s...->r...", np.asarray(T), u)
    yv = np.stack([resize(fftc(v[r], d, inverse=True, centered=False), list(shape)) for r in range(L * K)])
    yv = yv.reshape((L, K) + tuple(shape))
    return np.einsum("l...,lk...->k...", np.conj(maps), yv)


# --------------------------------------------------------------------------
# Cartesian masked FFT: src/pyfourier/_fft/_fft.py:101-151 with the B0 loops
# of :10-98 (each segment once, defect A.13 fixed).  x (B, F|K, *N) ->
# y (B, F, *N), all in centred order.  Forward: FFTc, then _basis_adj
# (y_t = sum_k conj(basis[t, k]) X_k, :141-151), then the mask; with
# b0 = (Bt (L, Nx), Ct (L, *N)): y = sum_l Bt[l, kx] * fwd(Ct_l * x), the
# temporal weight running along the readout (last) axis (_fft/_plan.py:83-96).
# --------------------------------------------------------------------------
def _mask_b(mask, ndim, F):
    if mask is None:
        return None
    m = np.asarray(mask, dtype=np.float64)
    m = m.reshape((-1,) + m.shape[-ndim:])
    if m.shape[0] not in (1, F):
        raise ValueError("mask frames must be 1 or F")
    return m[None]


def cart_fft(x, ndim, mask=None, basis=None, norm=None, b0=None):
    """Masked centred FFT.  x (B, F|K, *N) -> y (B, F, *N)."""
    x = np.asarray(x, dtype=np.complex128)
    basis = _basis_c(basis)
    F = basis.shape[0] if basis is not None else x.shape[1]
    segs = [(None, None)] if b0 is None else list(zip(b0[0], b0[1]))
    y = 0.0
    for bt, ct in segs:
        xs = x if ct is None else x * ct
        k = fftc(xs, ndim, norm=norm)
        if basis is not None:
            k = np.einsum("tk,bk...->bt...", np.conj(basis), k)
        if bt is not None:
            k = k * np.asarray(bt)
        y = y + k
    m = _mask_b(mask, ndim, F)
    return y if m is None else y * m


def cart_ifft(y, ndim, mask=None, basis=None, norm=None, b0=None):
    """Inverse of the above direction (_do_fft_adj, :116-138).  y (B, F, *N) -> x (B, F|K, *N)."""
    y = np.asarray(y, dtype=np.complex128)
    basis = _basis_c(basis)
    m = _mask_b(mask, ndim, y.shape[1])
    if m is not None:
        y = y * m
    segs = [(None, None)] if b0 is None else list(zip(b0[0], b0[1]))
    x = 0.0
    for bt, ct in segs:
        ys = y if bt is None else y * np.conj(np.asarray(bt))
        if basis is not None:
            ys = np.einsum("tk,bt...->bk...", basis, ys)
        xs = fftc(ys, ndim, norm=norm, inverse=True)
        if ct is not None:
            xs = xs * np.conj(ct)
        x = x + xs
    return x


def cart_gram_kernel(shape, nframes=1, mask=None, basis=None, bt=None):
    """Per-voxel Gram matrices in centred order, (nfr, *N, R, R) with
    R = L*K: S[(l',k'), (l,k)] = sum_t m_t conj(b_l') b_l basis[t,k'] conj(basis[t,k])
    (the intent of _gram/_plan.py:124-160: the mask once, outer products of
    the basis over contrasts; b_l = Bt[l, x] along the readout)."""
    d = len(shape)
    basis = _basis_c(basis)
    L = 1 if bt is None else np.asarray(bt).shape[0]
    K = 1 if basis is None else basis.shape[1]
    F = nframes if basis is None else basis.shape[0]
    m = np.ones((1,) + tuple(shape)) if mask is None else _mask_b(mask, d, F)[0]
    b = np.ones((L,) + (1,) * (d - 1) + (shape[-1],), dtype=np.complex128) if bt is None else \
        np.asarray(bt, dtype=np.complex128).reshape((L,) + (1,) * (d - 1) + (shape[-1],))
    bb = np.conj(b)[:, None] * b[None, :]  # (L', L, 1.., Nx)
    if basis is None:
        nfr = m.shape[0]
        S = m[:, None, None] * bb[None]  # (nfr, L', L, *N)
        S = np.moveaxis(S, (1, 2), (-2, -1))
        return S
    mt = np.broadcast_to(m, (F,) + tuple(shape))
    BB = basis[:, :, None] * np.conj(basis)[:, None, :]  # (t, k', k)
    Sk = np.einsum("t...,tpq->...pq", mt, BB)  # (*N, K', K)
    S = np.einsum("ab...,...pq->...apbq", bb, Sk).reshape(tuple(shape) + (L * K, L * K))
    return S[None]


def cart_gram(x, shape, nframes=1, mask=None, basis=None, b0=None):
    """Dense Gram (_gram/_gram.py:11-133, intent): per segment FFTc(Ct_l x_k),
    the per-voxel matrix S, IFFTc, sum conj(Ct_l'); both FFTs orthonormal, so
    the result equals ifftn(fftn(x)) for a binary mask under either norm."""
    x = np.asarray(x, dtype=np.complex128)
    d = len(shape)
    bt, ct = (None, None) if b0 is None else b0
    S = cart_gram_kernel(shape, nframes, mask, basis, bt)
    L = 1 if ct is None else len(ct)
    B, P = x.shape[:2]
    K = 1 if basis is None else np.asarray(basis).shape[1]
    u = np.stack([fftc(x if ct is None else x * ct[l], d, norm="ortho") for l in range(L)], axis=2)  # (B, P, L, *N)
    if basis is None:
        # (B, F, L, *N): per-frame L x L
        Sf = np.broadcast_to(S, (P,) + S.shape[1:]) if S.shape[0] == 1 else S
        v = np.einsum("f...pq,bfq...->bfp...", Sf, u)
    else:
        uk = u.transpose((0, 2, 1) + tuple(range(3, 3 + d))).reshape((B, L * K) + tuple(shape))
        v = np.einsum("...pq,bq...->bp...", S[0], uk).reshape((B, L, K) + tuple(shape)).transpose(
            (0, 2, 1) + tuple(range(3, 3 + d)))
    out = 0.0
    for l in range(L):
        xs = fftc(v[:, :, l], d, norm="ortho", inverse=True)
        if ct is not None:
            xs = xs * np.conj(ct[l])
        out = out + xs
    return out


# --------------------------------------------------------------------------
# Sparse FFT: centred FFT, then the sampled points picked
# (_dense2sparse/_2D.py:13-46: y[f, b, j] = X[f, b, idx_f[j]]; with a basis
# y[f] = sum_k basis[f, k] X_k[idx_f], _2D_subspace.py:13-54); the inverse
# zero-fills with repeated points summed and conj(basis), then inverse FFT.
# indexes (F, M, d) integer grid positions in [0, N), last axis (x, y, z).
# b0 = (Bt (L, nsamples), Ct (L, *N)) with the readout index j % nsamples.
# --------------------------------------------------------------------------
def _flat_index(indexes, shape):
    idx = np.asarray(indexes, dtype=np.int64)
    d = len(shape)
    flat = np.zeros(idx.shape[:-1], dtype=np.int64)
    for a in range(d):  # coordinate axis a (x, y, z) is grid axis d-1-a
        flat = flat * shape[a] + idx[..., d - 1 - a]
    return flat


def sparse_fft(x, shape, indexes, basis=None, weight=None, norm=None, b0=None):
    """x (B, F|K, *N) -> y (B, F, M)."""
    x = np.asarray(x, dtype=np.complex128)
    d = len(shape)
    flat = _flat_index(indexes, shape)  # (F, M)
    F, M = flat.shape
    basis = _basis_c(basis)
    B = x.shape[0]
    segs = [(None, None)] if b0 is None else list(zip(b0[0], b0[1]))
    y = np.zeros((B, F, M), dtype=np.complex128)
    for bt, ct in segs:
        xs = x if ct is None else x * ct
        k = fftc(xs, d, norm=norm).reshape(B, xs.shape[1], -1)
        if basis is None:
            ys = np.stack([k[:, f if k.shape[1] > 1 else 0, flat[f]] for f in range(F)], axis=1)
        else:
            ys = np.stack([sum(basis[f, q] * k[:, q, flat[f]] for q in range(basis.shape[1])) for f in range(F)],
                          axis=1)
        if bt is not None:
            t = np.asarray(bt).reshape(-1)
            ys = ys * np.tile(t, M // t.size)[None, None]
        y += ys
    if weight is not None:
        y = y * np.asarray(weight).reshape(1, F, -1)
    return y


def sparse_ifft(y, shape, indexes, basis=None, weight=None, norm=None, b0=None):
    """y (B, F, M) -> x (B, F|K, *N)."""
    y = np.asarray(y, dtype=np.complex128)
    d = len(shape)
    flat = _flat_index(indexes, shape)
    F, M = flat.shape
    B = y.shape[0]
    if weight is not None:
        y = y * np.asarray(weight).reshape(1, F, -1)
    basis = _basis_c(basis)
    P = F if basis is None else basis.shape[1]
    ntot = int(np.prod(shape))
    segs = [(None, None)] if b0 is None else list(zip(b0[0], b0[1]))
    x = 0.0
    for bt, ct in segs:
        ys = y if bt is None else y * np.conj(np.tile(np.asarray(bt).reshape(-1), M // np.asarray(bt).size))[None, None]
        g = np.zeros((B, P, ntot), dtype=np.complex128)
        for f in range(F):
            if basis is None:
                np.add.at(g[:, f], (slice(None), flat[f]), ys[:, f])
            else:
                for q in range(P):
                    np.add.at(g[:, q], (slice(None), flat[f]), np.conj(basis[f, q]) * ys[:, f])
        xs = fftc(g.reshape((B, P) + tuple(shape)), d, norm=norm, inverse=True)
        if ct is not None:
            xs = xs * np.conj(ct)
        x = x + xs
    return x


# --------------------------------------------------------------------------
# GROG: operators trained by regularised least squares on shifted
# calibration patches (:273-340), fractional powers by eigendecomposition
# (:344-367), the product table G[(sz, sy, sx)] = Gx^dx Gy^dy Gz^dz
# (:140-174) indexed by the quantised shift of each sample (:176-184), and
# every sample's coil row vector times its operator (the intent of
# :354-425).  Indexes round(c) + N//2 (x, y, z), weights 1/count of the
# flattened index (:188-218), samples off the matrix zeroed and their
# indexes clamped (:227-236).
# --------------------------------------------------------------------------
def grog_train(calib, ndim, lamda):
    c = np.moveaxis(np.asarray(calib, dtype=np.complex128), 0, -1)
    if c.ndim == 3:
        c = c[None]
    nc = c.shape[-1]

    def op(src, dst):
        S = src.reshape(src.shape[0], -1, nc) if ndim == 2 else src.reshape(1, -1, nc)
        T = dst.reshape(S.shape)
        out = []
        for s_, t_ in zip(S, T):
            Sh = s_.conj().T
            lam0 = lamda * np.sqrt((np.abs(Sh) ** 2).sum()) / nc
            out.append(np.linalg.solve(Sh @ s_ + lam0 * np.eye(nc), Sh @ t_))
        return np.stack(out)

    gx = op(c[:, :, :-1], c[:, :, 1:])
    gy = op(c[:, :-1], c[:, 1:])
    if ndim == 2:
        return [gx, gy]
    return [gx, gy, op(c[:-1], c[1:])]


def grog_power(G, delta):
    w, V = np.linalg.eig(G)
    return V @ np.diag(w.astype(np.complex128) ** delta) @ np.linalg.inv(V)


def grog_lut(coord, nsteps):
    """Per-axis step indexes (float32 arithmetic like the reference)."""
    c = np.asarray(coord, dtype=np.float32)
    d = np.round(c).astype(np.float32) - c  # numpy rounds half to even
    return (np.floor(np.float32(nsteps - 1) * d).astype(np.int64) + nsteps // 2)


def grog_interp(x, calib, coord, shape, lamda=0.01, nsteps=11):
    """x (B, nc, M), coord (M, d) grid units (x, y, z), shape (z, y, x)."""
    x = np.asarray(x, dtype=np.complex128)
    coord = np.asarray(coord, dtype=np.float32).reshape(-1, np.asarray(coord).shape[-1])
    d = coord.shape[-1]
    cal = np.asarray(calib, dtype=np.complex128)
    cal = cal / np.linalg.norm(cal)
    nsteps = 2 * (nsteps // 2) + 1
    ops = grog_train(cal, d, lamda)
    deltas = (np.arange(nsteps) - (nsteps - 1) // 2) / (nsteps - 1)
    nob = ops[0].shape[0]
    pw = [[[grog_power(g[o], dl) for dl in deltas] for o in range(nob)] for g in ops]  # [axis][slice][step]
    steps = grog_lut(coord, nsteps)  # (M, d)
    B, nc, M = x.shape
    out = np.zeros_like(x)
    N = list(shape)[::-1]  # (x, y, z)
    idx = (np.round(coord).astype(np.int64) + np.asarray(N) // 2)
    for b in range(B):
        o = b if nob > 1 else 0
        for s_ in range(M):
            G = np.eye(nc, dtype=np.complex128)
            for a in range(d):
                G = G @ pw[a][o][steps[s_, a]]
            out[b, :, s_] = x[b, :, s_] @ G
    flat = np.zeros(M, dtype=np.int64)
    unfold = 1
    for a in range(d):
        flat += unfold * idx[:, a]
        unfold *= N[a]
    _, inv, cnt = np.unique(flat, return_inverse=True, return_counts=True)
    weights = 1.0 / cnt[inv].astype(np.float32)
    outside = np.zeros(M, dtype=bool)
    for a in range(d):
        outside |= (idx[:, a] < 0) | (idx[:, a] >= N[a])
        idx[:, a] = np.clip(idx[:, a], 0, N[a] - 1)
    out[:, :, outside] = 0
    return out, idx, weights
