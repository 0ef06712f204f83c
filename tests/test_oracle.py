"""Pins of the CPU oracle (no GPU).

Each check compares the restatement with something it does not share code
or formulas with: scipy's modified Bessel function, numerical Fourier
integrals of the kernel, direct DFT sums, the FP64 NUDFT, exact adjoint
identities and hand-computed small cases.
"""

import math

import numpy as np
import pytest
from scipy import integrate, special

from oracle import nudft as U
from oracle import restate as R
from pyfourier_amd import demo


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---- kernel ------------------------------------------------------------------
@pytest.mark.parametrize("beta", [2.0, 6.99666, 8.99615, 13.8551])
def test_kb_matches_scipy_i0(beta):
    x = np.linspace(-1, 1, 2001)
    exact = special.i0(beta * np.sqrt(1 - x * x))
    # Abramowitz & Stegun 9.8.1 / 9.8.2 relative accuracy is a few 1e-7
    assert np.max(np.abs(R.kb(x, beta) / exact - 1)) < 6e-7


def test_kb_special_values():
    b = 8.99615
    assert R.kb(1.0, b) == pytest.approx(1.0, abs=0)
    assert R.kb(-1.0, b) == pytest.approx(1.0, abs=0)
    assert R.kb(1.0000001, b) == 0.0
    assert R.kb(0.0, b) == pytest.approx(special.i0(b), rel=6e-7)


def test_beta_closed_form():
    # pi * sqrt((W/s*(s-1/2))^2 - 0.8); W=4, s=2 -> pi*sqrt(8.2)
    assert float(R.kb_beta([4], [2.0])[0]) == pytest.approx(math.pi * math.sqrt(8.2), rel=1e-6)
    assert float(R.kb_beta([6], [1.25])[0]) == pytest.approx(math.pi * math.sqrt((6 / 1.25 * 0.75) ** 2 - 0.8), rel=1e-6)


def test_os_shape():
    assert R.os_shape([128, 100, 33], [1.25, 1.25, 1.5]) == [160, 125, 50]


def test_kernel_scale_is_sum_of_half_open_taps():
    b = 8.99615
    per_axis = sum(special.i0(b * np.sqrt(1 - u * u)) for u in (-1.0, -0.5, 0.0, 0.5))
    assert R.kernel_scale([b, b], [4, 4]) == pytest.approx(per_axis ** 2, rel=2e-6)
    assert R.kernel_scale([b, b], [1, 4]) == pytest.approx(per_axis, rel=1e-6)


def test_interp_tables_brute_force():
    b = 8.99615
    c = np.array([[3.0], [0.5]], dtype=np.float32)  # grid of 8, width 4
    idx, val = R.interp_tables(c, [4], [b], [8])
    # c=3: x0=ceil(1)=1 -> taps 1..4 at offsets -2,-1,0,1
    assert idx[0][0].tolist() == [1, 2, 3, 4]
    np.testing.assert_allclose(val[0][0], special.i0(b * np.sqrt(1 - np.array([1, .25, 0, .25]))), rtol=6e-7)
    # c=0.5: x0=ceil(-1.5)=-1 -> taps -1..2, wrapped mod 8
    assert idx[0][1].tolist() == [7, 0, 1, 2]
    u = np.array([-1.5, -0.5, 0.5, 1.5]) / 2
    np.testing.assert_allclose(val[0][1], special.i0(b * np.sqrt(1 - u * u)), rtol=6e-7)


@pytest.mark.parametrize("W,s,n", [(4, 2.0, 64), (6, 2.0, 48), (4, 1.25, 40), (4, 1.6, 10), (4, 1.6, 100)])
def test_apodization_inverts_kernel_transform(W, s, n):
    """apod(i) * phi_hat((i - n//2)/G) is constant, with phi_hat the Fourier
    transform of the exact KB kernel computed by quadrature and G the grid
    length the FFT runs on (os_shape: sigma*n in float32; at sigma 1.6 and n
    10 or 100 a float64 product would round up to 17 or 161)."""
    beta = float(R.kb_beta([W], [s])[0])
    G = R.os_shape([n], [s])[0]
    if s == 1.6:
        assert G == round(s * n) < math.ceil(float(np.float32(s)) * n)

    def phi_hat(u):
        f = lambda t: special.i0(beta * math.sqrt(max(0.0, 1 - (2 * t / W) ** 2))) * math.cos(2 * math.pi * u * t)
        return 2 * integrate.quad(f, 0, W / 2, limit=200, epsabs=0, epsrel=1e-12)[0]

    a = R.apod_vector(n, s, W, beta)
    prod = np.array([a[i] * phi_hat((i - n // 2) / G) for i in range(n)])
    assert np.max(np.abs(prod / prod[n // 2] - 1)) < 1e-6


def test_resize_known_answer():
    x = np.arange(1, 5)
    np.testing.assert_array_equal(R.resize(x, [8]), [0, 0, 1, 2, 3, 4, 0, 0])
    np.testing.assert_array_equal(R.resize(np.arange(8), [4]), [2, 3, 4, 5])
    # the centre element n//2 lands on o//2
    np.testing.assert_array_equal(R.resize(np.arange(3), [6]), [0, 0, 0, 1, 2, 0])


@pytest.mark.parametrize("n", [6, 7])
def test_centered_fft_is_centered_dft(n):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = np.arange(n) - n // 2
    direct = np.array([np.sum(x * np.exp(-2j * np.pi * k * r / n)) for k in r])
    np.testing.assert_allclose(R.fftc(x, 1), direct, rtol=1e-12, atol=1e-12)


# ---- operators ---------------------------------------------------------------
@pytest.mark.parametrize("W,bound", [(4, 2e-3), (6, 5e-5)])
def test_nufft_2d_vs_nudft(W, bound):
    rng = np.random.default_rng(1)
    N = 32
    coord = demo.golden_angle_radial_2d(16, 64)
    x = demo.complex_gaussian(rng, (N, N))
    P = R.Plan(coord, N, width=W, oversamp=2.0)
    y = R.nufft(x[None, None], P)[0, 0]
    yt = U.nudft(x[None], P.coord.reshape(-1, 2), [N, N])[0]
    assert rel(y, yt) < bound
    k = demo.complex_gaussian(rng, y.shape)
    xa = R.nufft_adj(k[None, None], P)[0, 0] * np.prod(P.os_shape)
    xt = U.nudft_adj(k[None], P.coord.reshape(-1, 2), [N, N])[0]
    assert rel(xa, xt) < bound


def test_nufft_1d_3d_vs_nudft():
    rng = np.random.default_rng(2)
    c1 = np.sort(rng.uniform(-np.pi, np.pi, (200, 1))).astype(np.float32)
    c1[0] = -np.pi
    x1 = demo.complex_gaussian(rng, (40,))
    P1 = R.Plan(c1, 40, width=6, oversamp=2.0)
    assert rel(R.nufft(x1[None, None], P1)[0, 0], U.nudft(x1[None], P1.coord.reshape(-1, 1), [40])[0]) < 5e-5
    c3 = demo.golden_means_radial_3d(64, 32)
    x3 = demo.complex_gaussian(rng, (12, 12, 12))
    P3 = R.Plan(c3, 12, width=6, oversamp=2.0)
    y3 = R.nufft(x3[None, None], P3)[0, 0]
    assert rel(y3, U.nudft(x3[None], P3.coord.reshape(-1, 3), [12, 12, 12])[0]) < 1e-4


def test_cartesian_reduces_to_centered_fft():
    rng = np.random.default_rng(3)
    N = 16
    ky, kx = np.meshgrid(np.arange(N) - N // 2, np.arange(N) - N // 2, indexing="ij")
    coord = np.stack([kx, ky], -1).astype(np.float32)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    assert P.width == [1, 1] and P.oversamp == [1.0, 1.0]
    x = demo.complex_gaussian(rng, (N, N))
    y = R.nufft(x[None, None], P)[0, 0].reshape(N, N)
    ref = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x.astype(np.complex128))))
    np.testing.assert_allclose(y, ref, rtol=1e-10, atol=1e-10)


def test_adjointness_ortho():
    rng = np.random.default_rng(4)
    coord = demo.golden_angle_radial_2d(12, 40)
    P = R.Plan(coord, 20, width=4, oversamp=1.5)
    x = demo.complex_gaussian(rng, (1, 1, 20, 20))
    y = demo.complex_gaussian(rng, (1, 1, P.npts))
    lhs = np.vdot(R.nufft(x, P, norm="ortho"), y)
    rhs = np.vdot(x, R.nufft_adj(y, P, norm="ortho"))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_subspace_equals_frame_loop():
    rng = np.random.default_rng(5)
    F, K, N = 6, 3, 20
    coord = np.stack([demo.golden_angle_radial_2d(3, 32) for _ in range(F)])
    for f in range(F):
        th = 0.7 * f
        rot = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        coord[f] = coord[f] @ rot.T
    basis = demo.t2_subspace_basis(rng, F, K)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    X = demo.complex_gaussian(rng, (1, K, N, N))
    y = R.nufft(X, P, basis=basis)[0]
    for f in range(F):
        Pf = _replan(P.coord[f : f + 1], N)  # single-frame plan on the same grid coordinates
        yf = sum(basis[f, k] * R.nufft(X[:, k : k + 1], Pf)[0, 0] for k in range(K))
        assert rel(y[f], yf) < 1e-12
        yt = U.nudft(np.tensordot(basis[f], X[0], axes=(0, 0))[None], P.coord[f].reshape(-1, 2), [N, N])[0]
        assert rel(y[f], yt) < 3e-3


def _replan(coord_grid, N):
    return R.Plan(coord_grid, N, width=4, oversamp=2.0, normalize=False)


def test_b0_model_vs_exact_field_map_signal():
    rng = np.random.default_rng(6)
    N = 24
    coord = demo.variable_density_spiral_2d(4, 256, N)
    zmap = demo.smooth_field_map(rng, (N, N), 100.0)
    t = np.linspace(0, 10e-3, 256)
    bt, ct = R.mri_exp_approx(zmap, t, L=8, nbins=40)
    # temporal interpolation of exp(-z t) over the map's own values
    z = 2j * math.pi * zmap.ravel()
    approx = bt.T @ ct.reshape(8, -1)  # (nt, npix)
    exact = np.exp(-np.outer(t, z))
    assert rel(approx, exact) < 5e-3
    P = R.Plan(coord, N, width=6, oversamp=2.0)
    x = demo.complex_gaussian(rng, (N, N))
    y = R.nufft(x[None, None], P, b0=(bt, ct))[0, 0]
    ye = U.nudft_b0(x[None], P.coord.reshape(-1, 2), (N, N), zmap, t)[0]
    assert rel(y, ye) < 1e-2


def test_toeplitz_exact_kernel_equals_gram_matrix():
    rng = np.random.default_rng(7)
    N = 12
    coord = demo.golden_angle_radial_2d(8, 24)
    P = R.Plan(coord, N, width=4, oversamp=2.0)
    cg = P.coord.reshape(-1, 2)
    w = rng.uniform(0.5, 1.5, cg.shape[0])
    T = R.toeplitz_kernel_exact(cg, (N, N), w)
    x = demo.complex_gaussian(rng, (N, N))
    gx = R.toeplitz_apply(x, T, (N, N))
    explicit = U.nudft_adj((w * U.nudft(x[None], cg, [N, N])[0])[None], cg, [N, N])[0]
    assert rel(gx, explicit) < 1e-10


@pytest.mark.parametrize("case", ["subspace", "b0", "subspace+b0"])
def test_toeplitz_block_equals_segmented_normal_operator(case):
    """The R x R block Toeplitz restatement with the exact PSF equals the
    normal operator A^H W A of the NUDFT-based subspace / segmented field-map
    model, built directly (no kernel):
      A x = { y_f = sum_l Bt[l, t] * NUDFT_f(Ct_l * sum_k basis[f, k] x_k) }
      A^H y = { sum_l conj(Ct_l) * sum_f conj(basis[f, k]) NUDFT_f^H(conj(Bt[l, t]) y_f) }.
    Complex basis and field-map factors, so a transposed block, a missing
    conjugate or a swapped (l, k) order fails."""
    rng = np.random.default_rng(17)
    N, F, K, L, ns = 10, 3, 2, 3, 40
    shape = (N, N)
    coord = np.stack([demo.golden_angle_radial_2d(2, ns)[0] @ np.array(
        [[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]]).T for a in (0.0, 0.9, 2.1)])[:, None]
    cg = R.Plan(coord, N, width=4, oversamp=2.0).coord.reshape(F, -1, 2).astype(np.float64)  # (F, 2*ns, 2)
    M = cg.shape[1]
    w = rng.uniform(0.5, 1.5, (F, M))
    use_b = case != "b0"
    use_z = case != "subspace"
    Kb = K if use_b else 1
    Lz = L if use_z else 1
    basis = (rng.standard_normal((F, Kb)) + 1j * rng.standard_normal((F, Kb))) if use_b else np.ones((F, 1))
    bt = (rng.standard_normal((Lz, M)) + 1j * rng.standard_normal((Lz, M))) if use_z else np.ones((1, M))
    ct = (rng.standard_normal((Lz,) + shape) + 1j * rng.standard_normal((Lz,) + shape)) if use_z else None
    maps = np.ones((1,) + shape) if ct is None else ct
    x = demo.complex_gaussian(rng, (Kb,) + shape).astype(np.complex128)

    out = np.zeros_like(x)
    for f in range(F):
        xf = np.tensordot(basis[f], x, axes=(0, 0))  # sum_k basis[f, k] x_k
        yf = sum(bt[l] * U.nudft((maps[l] * xf)[None], cg[f], list(shape))[0] for l in range(Lz))
        yf = w[f] * yf
        back = sum(np.conj(maps[l]) * U.nudft_adj((np.conj(bt[l]) * yf)[None], cg[f], list(shape))[0]
                   for l in range(Lz))
        out += np.conj(basis[f])[:, None, None] * back[None]

    frame = np.repeat(np.arange(F), M)
    tidx = np.tile(np.arange(M), F)
    coefs = (bt[:, tidx].T[:, :, None] * basis[frame][:, None, :]).reshape(F * M, Lz * Kb)
    T = R.toeplitz_kernel_block_exact(cg.reshape(-1, 2), shape, coefs, w.reshape(-1))
    gx = R.toeplitz_apply_block(x, T, shape, ct)
    assert rel(gx, out) < 1e-10


def test_toeplitz_block_reduces_to_plain():
    rng = np.random.default_rng(18)
    N = 12
    coord = demo.golden_angle_radial_2d(6, 24)
    cg = R.Plan(coord, N, width=4, oversamp=2.0).coord.reshape(-1, 2)
    w = rng.uniform(0.5, 1.5, cg.shape[0])
    x = demo.complex_gaussian(rng, (N, N))
    T1 = R.toeplitz_kernel_exact(cg, (N, N), w)
    Tb = R.toeplitz_kernel_block_exact(cg, (N, N), np.ones((cg.shape[0], 1)), w)
    assert rel(Tb[0, 0], T1) < 1e-12
    assert rel(R.toeplitz_apply_block(x[None], Tb, (N, N))[0], R.toeplitz_apply(x, T1, (N, N))) < 1e-12


def test_port_matches_numpy_restatement():
    from oracle import port as Pt

    rng = np.random.default_rng(8)
    coord = demo.golden_means_radial_3d(64, 32)
    P = R.Plan(coord, 16, width=4, oversamp=2.0)
    pp = Pt.PortPlan(P)
    x = demo.complex_gaussian(rng, (2, 1, 16, 16, 16))
    y0 = R.nufft(x, P)
    assert rel(Pt.nufft(x, pp), y0) < 5e-6
    assert rel(Pt.nufft_adj(y0.astype(np.complex64), pp), R.nufft_adj(y0, P)) < 5e-6


# ---- coordinate gradient -------------------------------------------------------
@pytest.mark.parametrize("beta", [2.0, 8.99615, 13.8551])
def test_kb_deriv_matches_finite_differences(beta):
    x = np.linspace(-0.999, 0.999, 1999)
    h = 1e-6
    fd = (R.kb(x + h, beta) - R.kb(x - h, beta)) / (2 * h)
    d = R.kb_deriv(x, beta)
    assert np.max(np.abs(fd - d)) < 1e-8 * np.max(np.abs(d))
    assert R.kb_deriv(1.5, beta) == 0.0


def _nudft_torch(x, k, shape):
    """Differentiable exact NUDFT (complex128): x (B, *N), k (M, d) -> (B, M)."""
    import torch

    d = len(shape)
    ph = 0
    for a in range(d):
        n = shape[a]
        r = torch.arange(n, dtype=torch.float64) - n // 2
        view = [1] * d
        view[a] = n
        ph = ph + k[:, d - 1 - a].reshape(-1, *([1] * d)) * r.reshape(view) / n
    return (torch.exp(-2j * math.pi * ph) * x[:, None]).flatten(2).sum(-1)


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_coord_grad_vs_autograd_nudft(shape):
    """Trajectory gradient of the restated NUFFT (W=6, sigma=2, accurate to
    ~1e-5) against torch autograd through the exact NUDFT, for the forward
    and the adjoint (norm None: the adjoint carries 1/G)."""
    import torch

    rng = np.random.default_rng(11)
    d, M, B = len(shape), 60, 2
    k = (rng.uniform(-0.5, 0.5, (M, d)) * np.array(shape[::-1])).astype(np.float32)
    P = R.Plan(k[None, None], shape, width=6, oversamp=2.0)
    x = demo.complex_gaussian(rng, (B,) + shape).astype(np.complex128)
    v = demo.complex_gaussian(rng, (B, M)).astype(np.complex128)
    # forward: L = Re <v, y>
    kt = torch.tensor(k.astype(np.float64), requires_grad=True)
    y = _nudft_torch(torch.tensor(x), kt, shape)
    y.retain_grad()
    (torch.tensor(v).conj() * y).real.sum().backward()
    got = R.coord_grad(x[:, None], P, y.grad.numpy()[:, None])
    assert rel(got, kt.grad.numpy()) < 3e-4
    # adjoint: x = A^H v / G, L = Re <u, x>
    G = float(np.prod(P.os_shape))
    kt = torch.tensor(k.astype(np.float64), requires_grad=True)
    npix = int(np.prod(shape))
    A = _nudft_torch(torch.eye(npix, dtype=torch.complex128).reshape((npix,) + shape), kt, shape)  # (npix, M)
    xa = ((A.conj() @ torch.tensor(v).T).T / G).reshape((B,) + shape)
    xa.retain_grad()
    u = demo.complex_gaussian(rng, (B,) + shape).astype(np.complex128)
    (torch.tensor(u).conj() * xa).real.sum().backward()
    got = R.coord_grad(xa.grad.numpy()[:, None], P, v[:, None], adjoint=True)
    assert rel(got, kt.grad.numpy()) < 3e-4


def test_coord_grad_vs_central_differences():
    """Exact derivative of the discrete forward operator against FP64
    central differences (step 1e-4 grid units) of the restated operator
    itself; a sample whose tap window would move within the step is excluded
    (the operator jumps there by KB(1)/scale)."""
    rng = np.random.default_rng(12)
    shape, M, B, h, W = (12, 10, 14), 40, 2, 1e-4, 4
    k = rng.uniform(-0.5, 0.5, (M, 3)) * np.array(shape[::-1])
    P = R.Plan(k[None, None], shape, width=W, oversamp=2.0, precise=True)
    x = demo.complex_gaussian(rng, (B, 1) + shape).astype(np.complex128)
    v = demo.complex_gaussian(rng, (B, 1, M)).astype(np.complex128)
    got = R.coord_grad(x, P, v)
    fd = np.zeros((M, 3))
    for a in range(3):
        kp, km = k.copy(), k.copy()
        kp[:, a] += h
        km[:, a] -= h
        # samples are independent: shifting every sample at once gives each
        # sample's own derivative
        yp = R.nufft(x, R.Plan(kp[None, None], shape, width=W, oversamp=2.0, precise=True))
        ym = R.nufft(x, R.Plan(km[None, None], shape, width=W, oversamp=2.0, precise=True))
        fd[:, a] = np.real(np.conj(v) * (yp - ym)).sum(axis=0).reshape(-1) / (2 * h)
    c = P.coord_os.reshape(-1, 3)
    step = h * np.array(P.os_shape[::-1]) / np.array(shape[::-1])
    ok = np.all(np.ceil(c + step - W / 2) == np.ceil(c - step - W / 2), axis=1)
    assert ok.sum() >= M - 3
    assert rel(got[ok], fd[ok]) < 1e-7


def test_coord_grad_adjoint_vs_central_differences():
    """Adjoint operator: d/dk_j of L = Re<u, nufft_adj(v)> by FP64 central
    differences, one sample moved at a time."""
    rng = np.random.default_rng(13)
    shape, M, B, h, W = (10, 12), 24, 2, 1e-4, 5
    k = rng.uniform(-0.5, 0.5, (M, 2)) * np.array(shape[::-1])
    P = R.Plan(k[None, None], shape, width=W, oversamp=1.5, precise=True)
    u = demo.complex_gaussian(rng, (B, 1) + shape).astype(np.complex128)
    v = demo.complex_gaussian(rng, (B, 1, M)).astype(np.complex128)
    got = R.coord_grad(u, P, v, adjoint=True)

    def loss(kk):
        Pk = R.Plan(kk[None, None], shape, width=W, oversamp=1.5, precise=True)
        return float(np.real(np.sum(np.conj(u) * R.nufft_adj(v, Pk))))

    c = P.coord_os.reshape(-1, 2)
    step = h * np.array(P.os_shape[::-1]) / np.array(shape[::-1])
    checked = 0
    for j in range(0, M, 3):
        if np.any(np.ceil(c[j] + step - W / 2) != np.ceil(c[j] - step - W / 2)):
            continue
        for a in range(2):
            kp, km = k.copy(), k.copy()
            kp[j, a] += h
            km[j, a] -= h
            fd = (loss(kp) - loss(km)) / (2 * h)
            assert fd == pytest.approx(got[j, a], rel=1e-6, abs=1e-9 * np.abs(got).max())
            checked += 1
    assert checked >= 10


# ---- Pipe-Menon density compensation ---------------------------------------------
def test_pipe_menon_cartesian_fixed_point():
    """A full Cartesian grid makes the ortho NUFFT a unitary DFT: G G^H = I,
    so w = 1 is the fixed point."""
    g = np.arange(16) - 8
    cc = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(1, 256, 2).astype(np.float32)
    w = R.pipe_menon_dcf(cc, 16, niter=4)
    assert np.max(np.abs(w - 1.0)) < 1e-9


@pytest.mark.parametrize("niter", [1, 2])
def test_pipe_menon_vs_exact_operator(niter):
    """The iteration with the FP64 NUDFT in place of the NUFFT (ortho scaling
    1/sqrt(G) on each side), accurate to the width-6 kernel error."""
    N = 24
    coord = demo.golden_angle_radial_2d(20, 48)
    P = R.Plan(coord, N, width=6, oversamp=1.5)
    cg = P.coord.reshape(-1, 2)
    G = float(np.prod(P.os_shape))
    w = np.ones((1, cg.shape[0]))
    for _ in range(niter):
        x = U.nudft_adj(w.astype(np.complex128), cg, [N, N]).real / math.sqrt(G)
        y = U.nudft(x.astype(np.complex128), cg, [N, N]).real / math.sqrt(G)
        w = w / np.abs(y)
    got = R.pipe_menon_dcf(coord, N, niter=niter)
    assert rel(got.ravel(), w.ravel()) < 2e-3


# ---------------------------------------------------------------------------
# Cartesian masked FFT and dense Gram
# ---------------------------------------------------------------------------
def _cart_grid(shape):
    """Centred Cartesian k-space points, (prod N, d) ordered (x, y[, z]) with x fastest."""
    g = np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij")
    return np.stack([a.ravel() for a in g[::-1]], axis=1).astype(np.float64)


@pytest.mark.parametrize("shape", [(9,), (6, 7), (4, 5, 6)])
def test_cart_fft_is_centred_dft_and_inverse(shape):
    rng = np.random.default_rng(40)
    x = demo.complex_gaussian(rng, (2, 1) + shape).astype(np.complex128)
    y = R.cart_fft(x, len(shape))
    direct = U.nudft(x[:, 0], _cart_grid(shape), list(shape))
    np.testing.assert_allclose(y.reshape(2, -1), direct, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(R.cart_ifft(y, len(shape)), x, rtol=1e-12, atol=1e-12)


def test_cart_fft_basis_is_conjugated_in_forward():
    """_basis_adj multiplies by basis.conj().T (src/pyfourier/_fft/_fft.py:141-151):
    with basis = i * I the forward returns -i * FFTc(X_t), the inverse +i."""
    rng = np.random.default_rng(41)
    x = demo.complex_gaussian(rng, (1, 3, 8, 6)).astype(np.complex128)
    basis = 1j * np.eye(3)
    y = R.cart_fft(x, 2, basis=basis)
    np.testing.assert_allclose(y, -1j * R.fftc(x, 2), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(R.cart_ifft(y, 2, basis=basis), x, rtol=1e-12, atol=1e-12)


def test_cart_fft_mask_frames_and_adjointness():
    rng = np.random.default_rng(42)
    T, K, shape = 5, 2, (6, 8)
    mask = (rng.random((T,) + shape) < 0.4).astype(np.float32)
    basis = demo.t2_subspace_basis(rng, T, K) * np.exp(1j * rng.random((T, K)))
    x = demo.complex_gaussian(rng, (2, K) + shape)
    y = R.cart_fft(x, 2, mask=mask, basis=basis, norm="ortho")
    assert np.all(y[:, mask == 0] == 0)
    w = demo.complex_gaussian(rng, (2, T) + shape)
    lhs = np.vdot(y, w)
    rhs = np.vdot(x, R.cart_ifft(w, 2, mask=mask, basis=basis, norm="ortho"))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # plain frames with one shared mask: each frame is its own masked FFT
    xf = demo.complex_gaussian(rng, (1, T) + shape).astype(np.complex128)
    yf = R.cart_fft(xf, 2, mask=mask[:1])
    for t in range(T):
        np.testing.assert_allclose(yf[0, t], mask[0] * R.fftc(xf[0, t], 2), rtol=1e-12, atol=1e-12)


def test_cart_fft_b0_exact_for_piecewise_constant_field_map():
    """With a field map of L distinct values, Ct_l = indicator of value l and
    Bt[l, t] = exp(-2 pi i f_l t) make the segmented model exact; it must then
    equal the direct field-map signal with t running along the readout (x)."""
    rng = np.random.default_rng(43)
    shape = (6, 10)
    freqs = np.array([-60.0, 0.0, 35.0, 90.0])
    labels = rng.integers(0, len(freqs), size=shape)
    zmap = freqs[labels]
    t = np.linspace(0, 4e-3, shape[-1])
    ct = np.stack([(labels == l).astype(np.complex128) for l in range(len(freqs))])
    bt = np.exp(-2j * math.pi * freqs[:, None] * t[None, :])
    x = demo.complex_gaussian(rng, (2, 1) + shape)
    y = R.cart_fft(x, 2, b0=(bt, ct))
    ye = U.nudft_b0(x[:, 0], _cart_grid(shape), shape, zmap, t)
    np.testing.assert_allclose(y.reshape(2, -1), ye, rtol=1e-10, atol=1e-10)
    # the inverse direction is the adjoint under ortho
    w = demo.complex_gaussian(rng, (2, 1) + shape)
    lhs = np.vdot(R.cart_fft(x, 2, b0=(bt, ct), norm="ortho"), w)
    rhs = np.vdot(x, R.cart_ifft(w, 2, b0=(bt, ct), norm="ortho"))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("case", ["plain", "frames", "basis", "b0", "basis_b0"])
def test_cart_gram_equals_inverse_of_forward(case):
    """The kernel form (per-voxel matrices between two FFTs) equals
    ifftn(fftn(x)) for a binary mask, whatever the norm."""
    rng = np.random.default_rng(44)
    shape = (8, 6)
    T, K, L = 4, 3, 2
    mask = (rng.random((T,) + shape) < 0.5).astype(np.float32)
    basis = b0 = None
    nimg = 1
    if case == "plain":
        mask = mask[:1]
    if case == "frames":
        nimg = T
    if case in ("basis", "basis_b0"):
        basis = demo.t2_subspace_basis(rng, T, K) * np.exp(1j * rng.random((T, K)))
        nimg = K
    if case in ("b0", "basis_b0"):
        if case == "b0":
            mask = mask[:1]
        bt = np.exp(-2j * math.pi * rng.random((L, shape[-1])))
        ct = np.exp(-2j * math.pi * rng.random((L,) + shape)) * rng.random((L,) + shape)
        b0 = (bt, ct)
    x = demo.complex_gaussian(rng, (2, nimg) + shape)
    g = R.cart_gram(x, shape, nframes=nimg if basis is None else 1, mask=mask, basis=basis, b0=b0)
    for norm in (None, "ortho"):
        ref = R.cart_ifft(R.cart_fft(x, 2, mask=mask, basis=basis, norm=norm, b0=b0), 2, mask=mask, basis=basis,
                          norm=norm, b0=b0)
        assert rel(g, ref) < 1e-12


# ---------------------------------------------------------------------------
# sparse FFT and GROG
# ---------------------------------------------------------------------------
def test_sparse_fft_full_grid_is_centred_dft_and_inverse():
    rng = np.random.default_rng(50)
    shape = (6, 8)
    g = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    idx = np.stack([g[1].ravel(), g[0].ravel()], -1)[None]  # (1, M, 2) x, y; every grid point once
    x = demo.complex_gaussian(rng, (2, 1) + shape).astype(np.complex128)
    y = R.sparse_fft(x, shape, idx)
    np.testing.assert_allclose(y.reshape((2, 1) + shape), R.fftc(x, 2), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(R.sparse_ifft(y, shape, idx), x, rtol=1e-12, atol=1e-12)


def test_sparse_fft_duplicates_sum_and_adjoint():
    rng = np.random.default_rng(51)
    shape = (5, 7)
    F, K = 3, 2
    idx = np.stack([rng.integers(0, 7, (F, 40)), rng.integers(0, 5, (F, 40))], -1)
    idx[:, 1] = idx[:, 0]  # a repeated point in every frame
    basis = demo.t2_subspace_basis(rng, F, K) * np.exp(1j * rng.random((F, K)))
    x = demo.complex_gaussian(rng, (2, K) + shape)
    w = demo.complex_gaussian(rng, (2, F, 40))
    y = R.sparse_fft(x, shape, idx, basis=basis, norm="ortho")
    lhs = np.vdot(y, w)
    rhs = np.vdot(x, R.sparse_ifft(w, shape, idx, basis=basis, norm="ortho"))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # the repeated point is gathered twice and zero-filled as the sum of both
    np.testing.assert_allclose(y[:, :, 0], y[:, :, 1], rtol=1e-12, atol=1e-12)
    e = np.zeros((1, 1, 40), dtype=np.complex128)
    e[0, 0, 0] = e[0, 0, 1] = 1.0
    g = R.sparse_ifft(e, shape, idx[:1], norm="ortho")
    k = R.fftc(g, 2, norm="ortho")
    iy, ix = idx[0, 0, 1], idx[0, 0, 0]
    assert abs(k[0, 0, iy, ix] - 2.0) < 1e-12


def _shift_model(rng, nc, ndim):
    """Commuting shift operators A_a = P diag(exp(i th_a)) P^-1 and calibration
    data generated exactly by them: c[..., y, x] = v @ Ay^y @ Ax^x."""
    P = np.eye(nc) + 0.3 * (rng.standard_normal((nc, nc)) + 1j * rng.standard_normal((nc, nc)))
    Pi = np.linalg.inv(P)
    th = [rng.uniform(-0.6, 0.6, nc) for _ in range(ndim)]
    A = [P @ np.diag(np.exp(1j * t)) @ Pi for t in th]

    def pw(a, e):
        return P @ np.diag(np.exp(1j * th[a] * e)) @ Pi

    return A, pw


def test_grog_training_recovers_exact_shift_operators():
    rng = np.random.default_rng(52)
    nc, n = 6, 10
    A, pw = _shift_model(rng, nc, 2)
    v = rng.standard_normal((3, nc)) + 1j * rng.standard_normal((3, nc))  # several sources for full rank
    cal = np.zeros((nc, 3, n, n), dtype=np.complex128)  # (nc, slices, y, x)
    for z in range(3):
        for yy in range(n):
            for xx in range(n):
                cal[:, z, yy, xx] = v[z] @ pw(1, yy) @ pw(0, xx)
    gx, gy = R.grog_train(cal, 2, 0.0)
    for z in range(3):
        assert rel(gx[z], A[0]) < 1e-8
        assert rel(gy[z], A[1]) < 1e-8


def test_grog_fractional_power_vs_schur_pade():
    from scipy.linalg import fractional_matrix_power

    rng = np.random.default_rng(53)
    A, _ = _shift_model(rng, 5, 1)
    for dl in (-0.5, -0.2, 0.1, 0.4):
        np.testing.assert_allclose(R.grog_power(A[0], dl), fractional_matrix_power(A[0], dl), rtol=1e-9,
                                   atol=1e-9)


def test_grog_lut_known_answers():
    # shift d = round(c) - c, step = floor(10 d) + 5 (nsteps 11); round half to even
    c = np.array([[0.26], [-1.24], [2.5], [3.5], [0.0], [-0.05]], dtype=np.float32)
    np.testing.assert_array_equal(R.grog_lut(c, 11)[:, 0], [2, 7, 0, 10, 5, 5])


def test_grog_moves_model_samples_by_the_quantised_shift():
    """On data generated by the shift model, a sample at c comes out as the
    model value at c + delta(step): the row-vector product with the x
    operator for the x shift and the y operator for the y shift."""
    from scipy.linalg import fractional_matrix_power

    rng = np.random.default_rng(54)
    nc, n = 4, 12
    A, pw = _shift_model(rng, nc, 2)
    v = rng.standard_normal((4, nc)) + 1j * rng.standard_normal((4, nc))
    cal = np.zeros((nc, n, n), dtype=np.complex128)
    for yy in range(n):
        for xx in range(n):
            cal[:, yy, xx] = sum(v[q] @ pw(1, yy + 0.37 * q) @ pw(0, xx - 0.21 * q) for q in range(4))
    M = 30
    coord = rng.uniform(-5.4, 5.4, (M, 2)).astype(np.float32)
    u = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    x = np.stack([u @ fractional_matrix_power(A[0], float(cx)) @ fractional_matrix_power(A[1], float(cy))
                  for cx, cy in coord.astype(np.float64)], -1)[None]  # (1, nc, M)
    out, idx, w = R.grog_interp(x, cal, coord, (n, n), lamda=0.0, nsteps=11)
    steps = R.grog_lut(coord, 11)
    delta = (steps - 5) / 10.0
    want = np.stack([u @ fractional_matrix_power(A[0], float(coord[s, 0]) + delta[s, 0])
                     @ fractional_matrix_power(A[1], float(coord[s, 1]) + delta[s, 1]) for s in range(M)], -1)
    assert rel(out[0], want) < 1e-6
    np.testing.assert_array_equal(idx, np.round(coord).astype(np.int64) + n // 2)


def test_grog_indexes_weights_and_outside_samples():
    rng = np.random.default_rng(55)
    nc, n = 3, 8
    cal = demo.complex_gaussian(rng, (nc, n, n))
    coord = np.array([[0.1, 0.2], [-0.2, 0.1], [3.6, 0.0], [-4.4, 1.0], [1.0, -1.0], [0.9, -1.2]], np.float32)
    x = demo.complex_gaussian(rng, (2, nc, 6))
    out, idx, w = R.grog_interp(x, cal, coord, (n, n))
    # samples 0, 1 share (4, 4); 4, 5 share (5, 3); 2 rounds to x = 8 (outside), 3 to x = 0 (inside).
    # The reference counts repeated points by the flattened index x + 8 y computed before clamping,
    # so the outside point (8, 4) and (0, 5) share the key 40 and count as a pair.
    np.testing.assert_array_equal(idx, [[4, 4], [4, 4], [7, 4], [0, 5], [5, 3], [5, 3]])
    np.testing.assert_allclose(w, [0.5, 0.5, 0.5, 0.5, 0.5, 0.5])
    assert np.all(out[:, :, 2] == 0) and np.all(out[:, :, 3] != 0)


def test_operator_restatement_rejects_wrong_rank():
    """y must be (B, F, M): a (B, M) array would be read with the wrong axes."""
    coord = demo.golden_angle_radial_2d(4, 16)
    P = R.Plan(coord, 8, width=4, oversamp=2.0)
    with pytest.raises(ValueError):
        R.nufft_adj(np.zeros((2, P.npts), np.complex64), P)
    with pytest.raises(ValueError):
        R.nufft(np.zeros((2, 8, 8), np.complex64), P)


# ---- Voronoi density compensation (oracle.voronoi) ------------------------

def test_voronoi_oracle_square_grid_closed_form():
    """6 x 6 grid of spacing 0.5: the 16 interior cells are 0.25, the 20 hull
    cells open.  The pipeline then gives, by hand: clipping at the 4th largest
    (round(0.118 * 36) = 4) leaves the areas; normalisation maps the open
    cells to 1 and the others to 0.25 times the hull value, and the k-space
    centre (the point at the origin, a hull point) scales that to 1."""
    from oracle import voronoi as V

    g = (np.stack(np.meshgrid(np.arange(6.0), np.arange(6.0)), -1).reshape(-1, 2) * 0.5).astype(np.float32)
    a = V.cell_areas(g).reshape(6, 6)
    interior = np.zeros((6, 6), bool)
    interior[1:-1, 1:-1] = True
    assert np.allclose(a[interior], 0.25, rtol=1e-12) and np.all(a[~interior] == 0)
    w = V.voronoi(g, nshots=1, fullspoke=True).reshape(6, 6)
    assert np.allclose(w[interior], 0.25, rtol=1e-6) and np.allclose(w[~interior], 1.0, rtol=1e-6)


def test_voronoi_oracle_hexagonal_lattice():
    """Interior cells of a triangular lattice of spacing a are regular
    hexagons of area sqrt(3)/2 a^2."""
    from oracle import voronoi as V

    a = 0.3
    pts = [((i + 0.5 * (j % 2)) * a, j * a * math.sqrt(3) / 2) for j in range(9) for i in range(9)]
    p = np.asarray(pts)
    ar = V.cell_areas(p).reshape(9, 9)
    assert np.allclose(ar[2:-2, 2:-2], math.sqrt(3) / 2 * a * a, rtol=1e-10)


def test_voronoi_oracle_duplicates_and_full_disk():
    """Copies of a point share its cell evenly; the full disk is the first
    shot rotated by 2 pi s / nshots (fullspoke) about the centre."""
    from oracle import voronoi as V

    g = (np.stack(np.meshgrid(np.arange(6.0), np.arange(6.0)), -1).reshape(-1, 2) * 0.5).astype(np.float32)
    d = np.concatenate([g, g[14:15]])  # one interior point twice
    w1 = V.voronoi(g, 1, True)
    w2 = V.voronoi(d, 1, True)
    # after normalisation the copies carry half the single point's relative weight
    r1 = w1[14] / w1[7]
    r2 = w2[14] / w2[7]
    assert abs(w2[14] - w2[-1]) < 1e-7 and abs(r2 - r1 / 2) < 1e-5 * r1
    # one spoke along +x, 4 full spokes: the second shot points along +y
    spoke = np.stack([np.linspace(-1, 1, 9), np.zeros(9)], -1)[None].astype(np.float32)
    w = V.voronoi(spoke, nshots=4, fullspoke=True)
    assert w.shape == (9,) and abs(w[4] - 0.25) < 1e-6  # k-space centre scaled to 1/nshots
