"""CPU tests of the C ABI boundary and the host-side logic (no kernel launch).

* the library loads and exports every entry point include/pfhip.h declares;
* host-side argument validation returns the documented status codes and a
  message before anything reaches the GPU;
* the Python host mirror (normalisation, plan parameters, field-map basis,
  device handling) matches the oracle's statement of the same rules.
"""

import ctypes
import math
import os
import re

import numpy as np
import pytest
import torch

import pyfourier_amd as pf
from oracle import restate as R
from pyfourier_amd import _lib, demo
from pyfourier_amd._plan import kb_beta, resolve_device

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pfhip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(pf_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported():
    names = _declared()
    assert len(names) >= 20
    h = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, f"not exported: {missing}"
    # the Python binding covers every declared entry point
    assert set(names) <= set(_lib.EXPORTED)


def test_version_and_error_string():
    L = _lib.lib()
    assert L.pf_version() == 100
    assert isinstance(L.pf_last_error(), bytes)


def _create(ndim, shape, width, oversamp, nframes=1, npts=16):
    L = _lib.lib()
    fake = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    sh = (ctypes.c_int * 3)(*shape)
    wi = (ctypes.c_int * 3)(*width)
    ov = (ctypes.c_float * 3)(*oversamp)
    out = ctypes.c_void_p()
    rc = L.pf_plan_create(fake, ndim, sh, wi, ov, nframes, npts, None, ctypes.byref(out))
    return rc, L.pf_last_error().decode(), out


@pytest.mark.parametrize(
    "args,code,msg",
    [
        ((4, (8, 8, 8), (4, 4, 4), (2.0, 2.0, 2.0)), 1, "ndim"),
        ((2, (8, 0, 0), (4, 4, 4), (2.0, 2.0, 2.0)), 1, "positive"),
        ((2, (8, 8, 0), (9, 4, 4), (2.0, 2.0, 2.0)), 5, "width"),
        ((2, (8, 8, 0), (4, 4, 4), (0.5, 2.0, 2.0)), 1, "oversampling"),
        ((1, (2, 0, 0), (4, 0, 0), (1.0, 0, 0)), 1, "smaller than the kernel"),
    ],
)
def test_plan_argument_validation(args, code, msg):
    rc, err, out = _create(*args)
    assert rc == code and msg in err and not out.value


def test_empty_trajectory_rejected():
    rc, err, _ = _create(2, (8, 8, 0), (4, 4, 0), (2.0, 2.0, 0), nframes=1, npts=0)
    assert rc == 1 and "empty" in err


def test_null_arguments():
    L = _lib.lib()
    d = _lib.BatchDesc()
    d.batch, d.nseg = 1, 1
    n = ctypes.c_size_t()
    assert L.pf_workspace_size(None, ctypes.byref(d), 0, ctypes.byref(n)) == 1
    assert L.pf_plan_create(None, 2, None, None, None, 1, 1, None, None) == 1
    assert L.pf_nufft(None, ctypes.byref(d), None, None, None, 0, None, 0, None) == 1


def test_device_resolution():
    with pytest.raises(RuntimeError, match="GPU only"):
        resolve_device("cpu")
    with pytest.raises(RuntimeError, match="GPU only"):
        resolve_device(-1)
    assert resolve_device("cuda:3") == torch.device("cuda", 3)
    assert resolve_device(2) == torch.device("cuda", 2)


@pytest.mark.parametrize(
    "maker,N",
    [
        (lambda: demo.golden_angle_radial_2d(64, 256), 128),
        (lambda: demo.golden_means_radial_3d(64, 32), (20, 24, 32)),
        (lambda: demo.variable_density_spiral_2d(8, 512, 48), 48),
        (lambda: demo.golden_angle_radial_2d(8, 16) / (2 * np.pi), 32),  # cycles/FOV units
        (lambda: demo.golden_angle_radial_2d(8, 16) * 3.0, 32),  # grid units (max |k| 3 pi <= N/2)
        (lambda: demo.golden_angle_radial_2d(8, 16) * 10.0, 32),  # other units (max |k| 10 pi > N/2)
    ],
)
def test_normalisation_matches_oracle(maker, N):
    c = maker().astype(np.float32)
    d = c.shape[-1]
    shape = [N] * d if np.isscalar(N) else list(N)
    mine = pf.normalize_coordinates(torch.from_numpy(c), shape[::-1]).numpy()
    np.testing.assert_array_equal(mine, R.normalize_coordinates(c, shape[::-1]))


# Closed forms of every unit branch (_subroutines/_traj.py:10-54 intent:
# any unit -> grid units [-N/2, N/2)), independent of both implementations.
def _norm(k, shape_xyz, **kw):
    return pf.normalize_coordinates(torch.from_numpy(np.asarray(k, np.float32)), shape_xyz, **kw).numpy()


def test_normalisation_radians():
    k = np.array([[-math.pi, 0.0], [0.5 * math.pi, 0.25]], dtype=np.float32)
    np.testing.assert_allclose(_norm(k, [64, 32]), k * np.array([64, 32]) / (2 * math.pi), rtol=1e-6)
    # a square k-space in radians: corners at radius pi*sqrt(2), edge at pi
    sq = np.array([[-math.pi, -math.pi], [math.pi * 0.99, math.pi * 0.5]], dtype=np.float32)
    np.testing.assert_allclose(_norm(sq, [64, 64]), sq * 64 / (2 * math.pi), rtol=1e-6)
    assert _norm(sq, [64, 64]).min() == pytest.approx(-32.0, rel=1e-6)


def test_normalisation_cycles():
    k = np.array([[-0.5, 0.125], [0.25, -0.375], [0.0, 0.4]], dtype=np.float32)
    # k * N: exact in float32 for these dyadic values
    np.testing.assert_array_equal(_norm(k, [64, 32]), k * np.array([64, 32], np.float32))
    # a square k-space in cycles/FOV (corners at radius 0.707 > 0.5)
    sq = np.array([[-0.5, -0.5], [0.25, 0.5]], dtype=np.float32)
    np.testing.assert_array_equal(_norm(sq, [16, 16]), sq * 16)


def test_normalisation_grid_units_pass_through():
    g = np.array([[-10.0, 3.0], [7.5, -20.0], [31.99, 32.0]], dtype=np.float32)
    np.testing.assert_array_equal(_norm(g, [64, 64]), g)
    # per-axis bound: |k_x| <= Nx/2 and |k_y| <= Ny/2 (the radius may exceed both)
    g2 = np.array([[-32.0, -8.0], [31.0, 7.9]], dtype=np.float32)
    np.testing.assert_array_equal(_norm(g2, [64, 16]), g2)
    # a small grid-unit trajectory (max |k| <= pi) is radians unless said otherwise
    small = np.array([[-3.0, 0.0], [1.5, 2.0]], dtype=np.float32)
    np.testing.assert_allclose(_norm(small, [64, 64]), small * 64 / (2 * math.pi), rtol=1e-6)
    np.testing.assert_array_equal(_norm(small, [64, 64], units="grid"), small)


def test_normalisation_physical_units():
    # 1/m: resolution 1 mm -> k_max = 500 /m; the edge maps to N/2
    kmax = 500.0
    k = np.array([[-kmax, 0.0], [0.5 * kmax, 0.25 * kmax], [0.1 * kmax, -kmax]], dtype=np.float32)
    out = _norm(k, [128, 96])
    np.testing.assert_allclose(out, k / kmax * np.array([64.0, 48.0]), rtol=1e-6)
    assert out[0, 0] == pytest.approx(-64.0, rel=1e-7) and out[2, 1] == pytest.approx(-48.0, rel=1e-7)
    # rad/m, 3-D, forced: same rule
    k3 = np.array([[0.0, 0.0, -3141.6], [1570.8, -1000.0, 200.0]], dtype=np.float32)
    np.testing.assert_allclose(_norm(k3, [32, 32, 16], units="max"), k3 / 3141.6 * np.array([16, 16, 8]), rtol=1e-6)


def test_normalisation_center_keyword_and_units_argument():
    k = np.array([[-0.5, 0.25]], dtype=np.float32)
    np.testing.assert_array_equal(_norm(k, [64, 32], center=False), k * np.array([64, 32], np.float32) + [32, 16])
    with pytest.raises(ValueError, match="units"):
        _norm(k, [64, 32], units="furlongs")


@pytest.mark.parametrize("bad", [float("nan"), float("inf"), -float("inf")])
def test_non_finite_coordinates_rejected(bad):
    c = demo.golden_angle_radial_2d(4, 16).astype(np.float32)
    c[1, 3, 0] = bad
    with pytest.raises(ValueError, match="finite"):
        pf.plan_nufft(torch.from_numpy(c), 32, width=4, oversamp=2.0)
    with pytest.raises(ValueError, match="finite"):
        pf.normalize_coordinates(torch.from_numpy(c), [32, 32])
    with pytest.raises(ValueError, match="finite"):
        R.normalize_coordinates(c, [32, 32])


def test_reference_import_surface():
    """The reference's top-level names, including the pipemenon module
    (src/pyfourier/__init__.py:9-15)."""
    for name in ("fftn", "ifftn", "plan_fft", "sparse_fftn", "sparse_ifftn", "plan_spfft", "nufft", "nufft_adj",
                 "plan_nufft", "gram", "plan_gram", "sparse_gram", "plan_spgram", "fft", "ifft"):
        assert callable(getattr(pf, name)), name
    assert callable(pf.pipemenon.pipe_menon_dcf)
    import pyfourier_amd.pipemenon as pm

    assert pm.pipe_menon_dcf is pf.pipe_menon_dcf


def test_gram_block_limit_checked_before_planning():
    """R = nseg * ncoeff above PF_TOEP_RMAX is refused before any PSF is
    computed (no GPU touched: the check precedes the native calls)."""
    c = demo.golden_angle_radial_2d(4, 16).astype(np.float32)
    basis = np.ones((1, _lib.PF_TOEP_RMAX + 1), np.complex64)
    with pytest.raises(ValueError, match="exceeds"):
        pf.plan_spgram(torch.from_numpy(c), 16, basis=basis, device="cuda:0")
    text = open(os.path.join(ROOT, "include", "pfhip.h")).read()
    assert f"#define PF_TOEP_RMAX {_lib.PF_TOEP_RMAX}" in text


def test_plan_parameters_on_host():
    coord = demo.golden_angle_radial_2d(16, 64)
    p = pf.plan_nufft(torch.from_numpy(coord), 100, width=4, oversamp=1.25)
    P = R.Plan(coord, 100, width=4, oversamp=1.25)
    assert list(p.os_shape) == P.os_shape == [125, 125]
    np.testing.assert_array_equal(np.asarray(p.beta, np.float32), np.asarray(P.beta, np.float32))
    assert p.interpolator.scale == pytest.approx(P.scale, rel=1e-12)
    np.testing.assert_array_equal(kb_beta([4, 6], [2.0, 1.25]), R.kb_beta([4, 6], [2.0, 1.25]))


def test_field_map_basis_host():
    rng = np.random.default_rng(0)
    zmap = demo.smooth_field_map(rng, (24, 24), 100.0)
    t = np.linspace(0, 10e-3, 200)
    b, ct = pf.mri_exp_approx(torch.from_numpy(zmap), torch.from_numpy(t), L=8, nbins=40)
    assert tuple(b.shape) == (8, 200) and tuple(ct.shape) == (8, 24, 24)
    bo, co = R.mri_exp_approx(zmap, t, L=8, nbins=40)
    assert np.linalg.norm(b.numpy() - bo) / np.linalg.norm(bo) < 1e-5
    assert np.linalg.norm(ct.numpy() - co) / np.linalg.norm(co) < 1e-5
    approx = b.numpy().astype(np.complex128).T @ ct.numpy().reshape(8, -1)
    exact = np.exp(-np.outer(t, 2j * math.pi * zmap.ravel()))
    assert np.linalg.norm(approx - exact) / np.linalg.norm(exact) < 5e-3


def test_spread_tile_listing_host_only():
    """pf_spread_tiles is host-only: it lists the register-spread tiles."""
    import pyfourier_amd as pf

    assert pf.spread_tiles(3)[0] == (8, 4, 4)
    assert all(len(t) == 2 for t in pf.spread_tiles(2))
    assert pf.spread_tiles(1) == [(64,)]
    L = _lib.lib()
    n = ctypes.c_int()
    assert L.pf_spread_tiles(4, None, 0, ctypes.byref(n)) == 1  # PF_ERR_ARG
