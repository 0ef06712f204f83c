// Cartesian operators (gfx950): masked centred FFT with subspace basis and
// time-segmented B0, the dense sibling of the NUFFT.
//
// Replaces the operator of src/pyfourier/_fft/_fft.py:10-151
//   forward  y_t = mask_t * sum_l b_l(x) * sum_k conj(basis[t, k]) * FFTc(Ct_l * X_k)
//   inverse  X_k = sum_l conj(Ct_l) * IFFTc(sum_t basis[t, k] * conj(b_l(x)) * mask_t * y_t)
// (_do_fft_fwd / _do_fft_adj at :101-126, _basis_fwd / _basis_adj at :129-151,
// the B0 segment loops at :10-98 with every segment added once, defect A.13
// fixed).  b_l is the temporal B0 interpolator along the readout (x) axis of
// centred k-space, src/pyfourier/_fft/_plan.py:83-96.
//
// Mask-only calls (no basis, no field map) on sizes of 2, 3 and 5 run the
// FFT passes directly on the caller's layout, centred on both sides, with the
// mask and norm factor fused into the last forward / first inverse pass and
// no workspace (pf_fftk_cart).  Otherwise the object's sample-free geometry
// plan (width 1, oversampling 1, unit apodisation) lets the fused FFT passes
// of the NUFFT (or rocFFT for sizes with other prime factors) compute FFTc on
// the (nfg, Cb, Nz, Ny, Nx, R, Cv) grid in stored FFT order.  One more kernel per direction moves between that grid and
// the caller's (C, T, Nz, Ny, Nx) centred k-space: it applies the shift, the
// basis contraction over K, the B0 sum over L, the mask and the norm factor in
// the same pass that transposes coils out of the innermost position.  Both are
// HBM-bound streaming kernels: a block stages a contiguous (x run, R, coil
// chunk) slab of the grid through LDS, so grid-side and k-space-side accesses
// are both unit-stride.
#include <algorithm>
#include <cmath>
#include <new>

#include "pf_internal.h"

struct pf_cart {
    pf_plan* geo = nullptr;   // Cartesian geometry: oversampling 1, width 1
    float* mask = nullptr;    // (mframes, Nz*Ny*Nx) centred order, or null
    int mframes = 0;
};

namespace {

constexpr int CT = 256;      // threads per block
constexpr int CXMAX = 256;   // x points per block (a whole row up to 256)
constexpr int CLDS = 8192;   // LDS budget per block in float2 (64 KB)
constexpr int CART_RMAX = 32;

struct CartGeo {
    int N[3];      // (z, y, x)
    int64_t ntot;  // Nz*Ny*Nx
};

struct CartBatch {
    int Cv, Cb, R, L, K, nfg, T, mframes;
    int cvb;   // coils per block (a divisor of Cv)
    int ncvb;  // Cv / cvb
    int cx;    // x points per block
    int nxc;   // x runs per row
    int tc;    // frames per block (sample kernel with one frame grid)
    int ntc;   // frame chunks
};

__device__ __forceinline__ int centred_of(int s, int N) {
    // stored FFT position s holds centred index (s + N//2) mod N
    int j = s + N / 2;
    return j >= N ? j - N : j;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

struct CartBlock {
    int fg, cb, x0, nx, cv0, ncv, zc, yc, t0, t1;
    int64_t zy;
};

__device__ __forceinline__ CartBlock cart_block(const CartGeo& g, const CartBatch& b) {
    CartBlock k;
    int64_t blk = blockIdx.x;
    const int tq = (int)(blk % b.ntc);
    blk /= b.ntc;
    const int xq = (int)(blk % b.nxc);
    blk /= b.nxc;
    const int cq = (int)(blk % b.ncvb);
    blk /= b.ncvb;
    const int64_t nzy = (int64_t)g.N[0] * g.N[1];
    k.zy = blk % nzy;
    blk /= nzy;
    k.cb = (int)(blk % b.Cb);
    k.fg = (int)(blk / b.Cb);
    k.x0 = xq * b.cx;
    k.nx = min(b.cx, g.N[2] - k.x0);
    k.cv0 = cq * b.cvb;
    k.ncv = b.cvb;
    k.zc = centred_of((int)(k.zy / g.N[1]), g.N[0]);
    k.yc = centred_of((int)(k.zy % g.N[1]), g.N[1]);
    // frames of this block: the frame grid's own frame, or a chunk of all
    k.t0 = b.nfg > 1 ? k.fg : tq * b.tc;
    k.t1 = b.nfg > 1 ? k.fg + 1 : min(b.T, k.t0 + b.tc);
    return k;
}

// grid -> centred k-space (forward)
template <int RMAX>
__global__ void __launch_bounds__(CT) k_cart_sample(CartGeo g, CartBatch b, const float2* __restrict__ grid,
                                                    float2* __restrict__ out, const float* __restrict__ mask,
                                                    const float2* __restrict__ basis, const float2* __restrict__ b0t,
                                                    float factor) {
    extern __shared__ float2 u[];
    const CartBlock k = cart_block(g, b);
    const int LD = b.R * b.cvb + 1;
    const int RC = b.R * b.Cv;
    const int Nx = g.N[2];
    const float2* src =
        grid + ((((int64_t)k.fg * b.Cb + k.cb) * ((int64_t)g.N[0] * g.N[1]) + k.zy) * Nx + k.x0) * RC;
    const int per_x = b.R * k.ncv;
    for (int q = threadIdx.x; q < k.nx * per_x; q += CT) {
        const int cvl = q % k.ncv;
        const int r = (q / k.ncv) % b.R;
        const int xl = q / per_x;
        u[xl * LD + r * b.cvb + cvl] = src[(int64_t)xl * RC + r * b.Cv + k.cv0 + cvl];
    }
    __syncthreads();
    const int t0 = k.t0, t1 = k.t1;
    for (int q = threadIdx.x; q < k.nx * k.ncv; q += CT) {
        const int xl = q % k.nx;
        const int cvl = q / k.nx;
        const int xc = centred_of(k.x0 + xl, Nx);
        const int64_t vox = ((int64_t)k.zc * g.N[1] + k.yc) * Nx + xc;
        const int64_t c = (int64_t)k.cb * b.Cv + k.cv0 + cvl;
        float2 v[RMAX];
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            v[r] = make_float2(0.f, 0.f);
            if (r < b.R) {
                v[r] = u[xl * LD + r * b.cvb + cvl];
                if (b0t) v[r] = cmul(v[r], b0t[(int64_t)(r / b.K) * Nx + xc]);
            }
        }
#pragma unroll 2
        for (int t = t0; t < t1; ++t) {
            float2 acc = make_float2(0.f, 0.f);
            if (basis) {
#pragma unroll
                for (int r = 0; r < RMAX; ++r)
                    if (r < b.R) {
                        const float2 w = basis[(int64_t)t * b.K + r % b.K];
                        acc = make_float2(acc.x + v[r].x * w.x + v[r].y * w.y, acc.y + v[r].y * w.x - v[r].x * w.y);
                    }
            } else {
#pragma unroll
                for (int r = 0; r < RMAX; ++r)
                    if (r < b.R) acc = make_float2(acc.x + v[r].x, acc.y + v[r].y);
            }
            float s = factor;
            if (mask) s *= mask[(int64_t)(b.mframes > 1 ? t : 0) * g.ntot + vox];
            out[(c * b.T + t) * g.ntot + vox] = make_float2(acc.x * s, acc.y * s);
        }
    }
}

// centred k-space -> grid (inverse direction, before the inverse FFT)
template <int RMAX>
__global__ void __launch_bounds__(CT) k_cart_scatter(CartGeo g, CartBatch b, const float2* __restrict__ in,
                                                     float2* __restrict__ grid, const float* __restrict__ mask,
                                                     const float2* __restrict__ basis,
                                                     const float2* __restrict__ b0t) {
    extern __shared__ float2 u[];
    const CartBlock k = cart_block(g, b);
    const int LD = b.R * b.cvb + 1;
    const int RC = b.R * b.Cv;
    const int Nx = g.N[2];
    const int t0 = k.t0, t1 = k.t1;
    for (int q = threadIdx.x; q < k.nx * k.ncv; q += CT) {
        const int xl = q % k.nx;
        const int cvl = q / k.nx;
        const int xc = centred_of(k.x0 + xl, Nx);
        const int64_t vox = ((int64_t)k.zc * g.N[1] + k.yc) * Nx + xc;
        const int64_t c = (int64_t)k.cb * b.Cv + k.cv0 + cvl;
        float2 acc[RMAX];
#pragma unroll
        for (int r = 0; r < RMAX; ++r) acc[r] = make_float2(0.f, 0.f);
#pragma unroll 4
        for (int t = t0; t < t1; ++t) {
            float2 y = in[(c * b.T + t) * g.ntot + vox];
            if (mask) {
                const float m = mask[(int64_t)(b.mframes > 1 ? t : 0) * g.ntot + vox];
                y = make_float2(y.x * m, y.y * m);
            }
            if (basis) {
#pragma unroll
                for (int r = 0; r < RMAX; ++r)
                    if (r < b.R) {
                        const float2 w = basis[(int64_t)t * b.K + r % b.K];
                        acc[r] = make_float2(acc[r].x + w.x * y.x - w.y * y.y, acc[r].y + w.x * y.y + w.y * y.x);
                    }
            } else {
#pragma unroll
                for (int r = 0; r < RMAX; ++r)
                    if (r < b.R) acc[r] = make_float2(acc[r].x + y.x, acc[r].y + y.y);
            }
        }
#pragma unroll
        for (int r = 0; r < RMAX; ++r)
            if (r < b.R) {
                float2 a = acc[r];
                if (b0t) a = cmulc(a, b0t[(int64_t)(r / b.K) * Nx + xc]);
                u[xl * LD + r * b.cvb + cvl] = a;
            }
    }
    __syncthreads();
    float2* dst = grid + ((((int64_t)k.fg * b.Cb + k.cb) * ((int64_t)g.N[0] * g.N[1]) + k.zy) * Nx + k.x0) * RC;
    const int per_x = b.R * k.ncv;
    for (int q = threadIdx.x; q < k.nx * per_x; q += CT) {
        const int cvl = q % k.ncv;
        const int r = (q / k.ncv) % b.R;
        const int xl = q / per_x;
        dst[(int64_t)xl * RC + r * b.Cv + k.cv0 + cvl] = u[xl * LD + r * b.cvb + cvl];
    }
}

size_t align_up(size_t n) { return (n + 255) / 256 * 256; }

int cart_batch(const pf_cart* c, const pf_batch_desc* d, int nframes, pf_batch* b) {
    if (!c) return pf_fail(PF_ERR_ARG, "null Cartesian plan");
    if (!d) return pf_fail(PF_ERR_ARG, "null batch description");
    if (d->batch < 1) return pf_fail(PF_ERR_ARG, "batch must be >= 1");
    if (nframes < 1) return pf_fail(PF_ERR_ARG, "nframes must be >= 1");
    if (d->nseg < 1) return pf_fail(PF_ERR_ARG, "nseg must be >= 1");
    if (d->ncoeff < 0) return pf_fail(PF_ERR_ARG, "ncoeff must be >= 0");
    if (d->ncoeff > 0 && !d->basis) return pf_fail(PF_ERR_ARG, "ncoeff > 0 needs a basis");
    if (c->mframes > 1 && c->mframes != nframes)
        return pf_fail(PF_ERR_ARG, "the mask has a frame count other than 1 and nframes");
    const bool has_b0 = d->b0_temporal || d->b0_spatial;
    if ((d->nseg > 1 || has_b0) && (!d->b0_temporal || !d->b0_spatial))
        return pf_fail(PF_ERR_ARG, "a field map needs both b0_temporal and b0_spatial");
    if (has_b0 && d->nt != c->geo->shape[2])
        return pf_fail(PF_ERR_ARG, "the B0 temporal interpolator must have one column per readout (x) point");
    b->C = d->batch;
    b->Cv = 1;
    for (int v = std::min(64, d->batch); v >= 1; --v)
        if (d->batch % v == 0) {
            b->Cv = v;
            break;
        }
    b->Cb = b->C / b->Cv;
    b->L = d->nseg;
    b->K = d->ncoeff > 0 ? d->ncoeff : 1;
    b->R = b->L * b->K;
    if (b->R > CART_RMAX) return pf_fail(PF_ERR_UNSUPPORTED, "nseg*ncoeff above 32 in the Cartesian operator");
    b->basis = d->ncoeff > 0 ? static_cast<const float2*>(d->basis) : nullptr;
    b->b0t = has_b0 ? static_cast<const float2*>(d->b0_temporal) : nullptr;
    b->b0s = has_b0 ? static_cast<const float2*>(d->b0_spatial) : nullptr;
    b->nt = has_b0 ? d->nt : 1;
    b->nfg = b->basis ? 1 : nframes;
    b->nimg = b->basis ? b->K : nframes;
    return PF_OK;
}

size_t cart_grid_bytes(const pf_cart* c, const pf_batch& b) {
    return (size_t)b.nfg * b.C * b.R * (size_t)pf_gtot(c->geo) * sizeof(float2);
}

// mask-only calls on sizes the fused FFT supports run the FFT passes directly
// on the caller's layout (no grid, no transposes, no workspace)
bool cart_direct(const pf_cart* c, const pf_batch& b) {
    return b.R == 1 && !b.basis && !b.b0t && pf_fftk_supported(c->geo);
}

size_t cart_ws_bytes(const pf_cart* c, const pf_batch& b) {
    if (cart_direct(c, b)) return 0;
    size_t extra = pf_fftk_supported(c->geo) ? pf_fftk_extra_bytes(c->geo, b) : 0;
    return align_up(cart_grid_bytes(c, b)) + align_up(extra);
}

CartBatch cart_kbatch(const pf_cart* c, const pf_batch& b, int nframes, bool fwd) {
    CartBatch k;
    k.Cv = b.Cv;
    k.Cb = b.Cb;
    k.R = b.R;
    k.L = b.L;
    k.K = b.K;
    k.nfg = b.nfg;
    k.T = nframes;
    k.mframes = c->mframes;
    // x run: a whole row up to 256 points; coils per block: the largest
    // divisor of Cv whose (x run, R, coil) slab fits the LDS budget
    const int Nx = c->geo->shape[2];
    k.cx = std::min(Nx, CXMAX);
    while (k.cx > 1 && k.cx * (b.R + 1) > CLDS) k.cx /= 2;
    const int lim = std::max(1, (CLDS / k.cx - 1) / b.R);
    k.cvb = 1;
    for (int v = std::min(b.Cv, lim); v >= 1; --v)
        if (b.Cv % v == 0) {
            k.cvb = v;
            break;
        }
    k.ncvb = b.Cv / k.cvb;
    k.nxc = (Nx + k.cx - 1) / k.cx;
    // forward with one frame grid (basis) and many frames: split the frames
    // over blocks until about 4096 blocks are in flight (each re-reads its
    // small grid slab); the adjoint sums over frames inside one block
    k.tc = nframes;
    k.ntc = 1;
    if (fwd && b.nfg == 1 && nframes > 1) {
        const int64_t base = (int64_t)b.Cb * c->geo->shape[0] * c->geo->shape[1] * k.ncvb * k.nxc;
        const int64_t want = std::min<int64_t>(nframes, (4096 + base - 1) / base);
        k.tc = (int)((nframes + want - 1) / want);
        k.ntc = (nframes + k.tc - 1) / k.tc;
    }
    return k;
}

template <int RMAX>
void launch_pair(bool fwd, unsigned nblk, size_t lds, hipStream_t s, const CartGeo& g, const CartBatch& k,
                 const float2* grid_in, float2* grid_out, const float2* ks_in, float2* ks_out, const float* mask,
                 const float2* basis, const float2* b0t, float factor) {
    if (fwd)
        hipLaunchKernelGGL(k_cart_sample<RMAX>, dim3(nblk), dim3(CT), lds, s, g, k, grid_in, ks_out, mask, basis,
                           b0t, factor);
    else
        hipLaunchKernelGGL(k_cart_scatter<RMAX>, dim3(nblk), dim3(CT), lds, s, g, k, ks_in, grid_out, mask, basis,
                           b0t);
}

int launch_cart(const pf_cart* c, const pf_batch& b, int nframes, bool fwd, const float2* grid_in, float2* grid_out,
                const float2* ks_in, float2* ks_out, float factor, hipStream_t s) {
    const pf_plan* p = c->geo;
    CartGeo g;
    for (int a = 0; a < 3; ++a) g.N[a] = p->shape[a];
    g.ntot = pf_ntot(p);
    CartBatch k = cart_kbatch(c, b, nframes, fwd);
    const int64_t nblk = (int64_t)b.nfg * b.Cb * p->shape[0] * p->shape[1] * k.ncvb * k.nxc * k.ntc;
    if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many Cartesian blocks");
    if (nblk < 1) return PF_OK;
    const size_t lds = (size_t)k.cx * (b.R * k.cvb + 1) * sizeof(float2);
    const unsigned nb = (unsigned)nblk;
    if (b.R == 1)
        launch_pair<1>(fwd, nb, lds, s, g, k, grid_in, grid_out, ks_in, ks_out, c->mask, b.basis, b.b0t, factor);
    else if (b.R <= 4)
        launch_pair<4>(fwd, nb, lds, s, g, k, grid_in, grid_out, ks_in, ks_out, c->mask, b.basis, b.b0t, factor);
    else if (b.R <= 8)
        launch_pair<8>(fwd, nb, lds, s, g, k, grid_in, grid_out, ks_in, ks_out, c->mask, b.basis, b.b0t, factor);
    else if (b.R <= 16)
        launch_pair<16>(fwd, nb, lds, s, g, k, grid_in, grid_out, ks_in, ks_out, c->mask, b.basis, b.b0t, factor);
    else
        launch_pair<32>(fwd, nb, lds, s, g, k, grid_in, grid_out, ks_in, ks_out, c->mask, b.basis, b.b0t, factor);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

// forward factor and inverse factor per norm code
void cart_factors(const pf_cart* c, int norm, double* fwd, double* inv) {
    const double n = (double)pf_gtot(c->geo);
    *fwd = 1.0;
    *inv = 1.0 / n;
    if (norm == PF_NORM_ORTHO) *fwd = *inv = 1.0 / std::sqrt(n);
    if (norm == PF_NORM_ADJOINT) {
        *fwd = 1.0 / n;
        *inv = 1.0;
    }
}

}  // namespace

extern "C" {

int pf_cart_create(const float* mask, int mframes, int ndim, const int* shape, void* stream, pf_cart** out) {
    if (!out) return pf_fail(PF_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!shape) return pf_fail(PF_ERR_ARG, "null shape");
    if (mask && mframes < 1) return pf_fail(PF_ERR_ARG, "a mask needs mframes >= 1");
    pf_cart* c = new (std::nothrow) pf_cart();
    if (!c) return pf_fail(PF_ERR_ALLOC, "host allocation failed");
    c->geo = new (std::nothrow) pf_plan();
    int w[3] = {1, 1, 1};
    float sg[3] = {1.0f, 1.0f, 1.0f};
    int rc = c->geo ? pf_plan_params(c->geo, ndim, shape, w, sg, 1, 1, nullptr, 0)
                    : pf_fail(PF_ERR_ALLOC, "host allocation failed");
    if (rc == PF_OK) rc = pf_plan_upload_apod(c->geo);
    if (rc == PF_OK) rc = pf_fftk_supported(c->geo) ? pf_fftk_init(c->geo) : pf_fft_cache_init(c->geo);
    if (rc == PF_OK && mask) {
        c->mframes = mframes;
        const size_t n = (size_t)mframes * pf_ntot(c->geo) * sizeof(float);
        hipError_t e = hipMalloc(&c->mask, n);
        if (e == hipSuccess) e = hipMemcpyAsync(c->mask, mask, n, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream));
        if (e == hipSuccess) e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
        if (e != hipSuccess) rc = pf_fail(PF_ERR_HIP, std::string("mask upload: ") + hipGetErrorString(e));
    }
    if (rc != PF_OK) {
        if (c->mask) (void)hipFree(c->mask);
        if (c->geo) pf_plan_free(c->geo);
        delete c;
        return rc;
    }
    *out = c;
    return PF_OK;
}

int pf_cart_destroy(pf_cart* c) {
    if (!c) return PF_OK;
    if (c->mask) (void)hipFree(c->mask);
    if (c->geo) pf_plan_free(c->geo);
    delete c;
    return PF_OK;
}

int pf_cart_workspace_size(const pf_cart* c, const pf_batch_desc* d, int nframes, size_t* bytes) {
    pf_batch b;
    int rc = cart_batch(c, d, nframes, &b);
    if (rc != PF_OK) return rc;
    if (!bytes) return pf_fail(PF_ERR_ARG, "null output");
    *bytes = cart_ws_bytes(c, b);
    return PF_OK;
}

int pf_cart_fft(pf_cart* c, const pf_batch_desc* d, int nframes, const void* image, void* kspace, int norm,
                void* workspace, size_t workspace_bytes, void* stream) {
    pf_batch b;
    int rc = cart_batch(c, d, nframes, &b);
    if (rc != PF_OK) return rc;
    if (!image || !kspace) return pf_fail(PF_ERR_ARG, "null array");
    if (norm < PF_NORM_NONE || norm > PF_NORM_ADJOINT) return pf_fail(PF_ERR_ARG, "norm must be 0, 1 or 2");
    if (workspace_bytes < cart_ws_bytes(c, b) || (!workspace && cart_ws_bytes(c, b) > 0))
        return pf_fail(PF_ERR_ARG, "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double fwd, inv;
    cart_factors(c, norm, &fwd, &inv);
    pf_plan* p = c->geo;
    if (cart_direct(c, b))
        return pf_fftk_cart(p, (int64_t)b.C * nframes, nframes, c->mask, c->mframes, true,
                            static_cast<const float2*>(image), static_cast<float2*>(kspace), (float)fwd, s);
    float2* grid = static_cast<float2*>(workspace);
    float2* extra = reinterpret_cast<float2*>(static_cast<char*>(workspace) + align_up(cart_grid_bytes(c, b)));
    if (pf_fftk_supported(p)) {
        if ((rc = pf_fftk_forward(p, b, static_cast<const float2*>(image), grid, extra, s)) != PF_OK) return rc;
    } else {
        if ((rc = pf_launch_deapod_pad(p, b, static_cast<const float2*>(image), grid, s)) != PF_OK) return rc;
        if ((rc = pf_fft_exec(p, b, grid, 0, s)) != PF_OK) return rc;
    }
    return launch_cart(c, b, nframes, true, grid, nullptr, nullptr, static_cast<float2*>(kspace), (float)fwd, s);
}

int pf_cart_ifft(pf_cart* c, const pf_batch_desc* d, int nframes, const void* kspace, void* image, int norm,
                 void* workspace, size_t workspace_bytes, void* stream) {
    pf_batch b;
    int rc = cart_batch(c, d, nframes, &b);
    if (rc != PF_OK) return rc;
    if (!image || !kspace) return pf_fail(PF_ERR_ARG, "null array");
    if (norm < PF_NORM_NONE || norm > PF_NORM_ADJOINT) return pf_fail(PF_ERR_ARG, "norm must be 0, 1 or 2");
    if (workspace_bytes < cart_ws_bytes(c, b) || (!workspace && cart_ws_bytes(c, b) > 0))
        return pf_fail(PF_ERR_ARG, "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double fwd, inv;
    cart_factors(c, norm, &fwd, &inv);
    pf_plan* p = c->geo;
    if (cart_direct(c, b))
        return pf_fftk_cart(p, (int64_t)b.C * nframes, nframes, c->mask, c->mframes, false,
                            static_cast<const float2*>(kspace), static_cast<float2*>(image), (float)inv, s);
    float2* grid = static_cast<float2*>(workspace);
    float2* extra = reinterpret_cast<float2*>(static_cast<char*>(workspace) + align_up(cart_grid_bytes(c, b)));
    if ((rc = launch_cart(c, b, nframes, false, nullptr, grid, static_cast<const float2*>(kspace), nullptr, 1.0f, s)) !=
        PF_OK)
        return rc;
    if (pf_fftk_supported(p))
        return pf_fftk_adjoint(p, b, grid, static_cast<float2*>(image), extra, (float)inv, s);
    if ((rc = pf_fft_exec(p, b, grid, 1, s)) != PF_OK) return rc;
    return pf_launch_crop_deapod(p, b, grid, static_cast<float2*>(image), (float)inv, s);
}

}  // extern "C"
