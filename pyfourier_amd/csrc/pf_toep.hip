// Toeplitz-embedded Gram operator (gfx950).
//
// Replaces the apply of src/pyfourier/_sparsegram/_sparsegram.py:11-81
// (pad to 2N -> FFT -> (x) kernel -> IFFT -> crop, with the B0 segment loop)
// and the per-voxel matvec of src/pyfourier/_subroutines/_bmatvec/_bdot.py:16-157.
//
// The object holds a sample-free geometry plan for the 2N embedding (width 1,
// oversampling 2, unit apodisation), so the pruned fused FFT passes of the
// NUFFT do the embedding: only rows that hold image data are transformed
// forward, and the inverse keeps only the N outputs the crop needs.  Between
// them one kernel applies the per-voxel R x R kernel (R = L segments x K
// subspace coefficients; R = 1 is a pointwise multiply).  With B0 segments the
// X pass multiplies by Ct_l on the way in and sums conj(Ct_l) on the way out.
//
// The same object, built on the N grid itself (oversampling 1, no embedding)
// and with one kernel per frame, is the dense Cartesian Gram of
// src/pyfourier/_gram/_gram.py:11-133 (pf_gram_create): there the kernel is
// the sampling mask (or its per-voxel subspace/B0 matrix) in FFT order.
#include <algorithm>
#include <new>
#include <string>

#include "pf_internal.h"

struct pf_toep {
    pf_plan* geo = nullptr;   // embedding geometry (no samples): 2N, or N for the dense Gram
    float2* T = nullptr;      // (nfr, Gtot, R, R) kernel in stored FFT order
    int R = 1;
    int L = 1;
    int K = 1;
    int nfr = 1;              // kernel frames: each image frame has its own kernel
};

namespace {

constexpr int TB = 256;

// grid (nfr, Cb, Gtot, R, Cv) in place: u[r] <- sum_r' T[f, pt, r, r'] u[r']
template <int RMAX>
__global__ void __launch_bounds__(TB) k_toep_mul(const float2* __restrict__ T, float2* __restrict__ grid,
                                                 int64_t gtot, int R, int Cv, int Cb, int nfr) {
    const int64_t total = (int64_t)nfr * Cb * gtot * Cv;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int cv = (int)(e % Cv);
        const int64_t pt = (e / Cv) % gtot;
        const int64_t fcb = e / (Cv * gtot);
        const int64_t f = fcb / Cb;
        float2* g = grid + (fcb * gtot + pt) * (int64_t)R * Cv + cv;
        const float2* t = T + (f * gtot + pt) * (int64_t)R * R;
        if (RMAX == 1) {
            const float2 a = t[0], u = g[0];
            g[0] = make_float2(a.x * u.x - a.y * u.y, a.x * u.y + a.y * u.x);
            continue;
        }
        float2 u[RMAX];
#pragma unroll
        for (int r = 0; r < RMAX; ++r)
            if (r < R) u[r] = g[r * Cv];
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            if (r >= R) break;
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll
            for (int q = 0; q < RMAX; ++q) {
                if (q >= R) break;
                const float2 a = t[r * R + q];
                acc.x += a.x * u[q].x - a.y * u[q].y;
                acc.y += a.x * u[q].y + a.y * u[q].x;
            }
            g[r * Cv] = acc;
        }
    }
}

// Blocks larger than any unrolled instance (16 < R <= PF_TOEP_RMAX): the same
// in-place product with each thread's input vector staged in LDS (column-major
// over the block's threads, so the q-th values of consecutive threads are
// consecutive words).
constexpr int TB_GEN = 64;
__global__ void __launch_bounds__(TB_GEN) k_toep_mul_lds(const float2* __restrict__ T, float2* __restrict__ grid,
                                                         int64_t gtot, int R, int Cv, int Cb, int nfr) {
    extern __shared__ float2 u_lds[];  // (R, TB_GEN)
    const int64_t total = (int64_t)nfr * Cb * gtot * Cv;
    const int64_t e = (int64_t)blockIdx.x * TB_GEN + threadIdx.x;
    if (e >= total) return;  // no block-wide barrier below: each thread uses only its own column
    const int cv = (int)(e % Cv);
    const int64_t pt = (e / Cv) % gtot;
    const int64_t fcb = e / (Cv * gtot);
    const int64_t f = fcb / Cb;
    float2* g = grid + (fcb * gtot + pt) * (int64_t)R * Cv + cv;
    const float2* t = T + (f * gtot + pt) * (int64_t)R * R;
    float2* u = u_lds + threadIdx.x;
    for (int q = 0; q < R; ++q) u[q * TB_GEN] = g[(int64_t)q * Cv];
    for (int r = 0; r < R; ++r) {
        float2 acc = make_float2(0.f, 0.f);
        for (int q = 0; q < R; ++q) {
            const float2 a = t[r * R + q], v = u[q * TB_GEN];
            acc.x += a.x * v.x - a.y * v.y;
            acc.y += a.x * v.y + a.y * v.x;
        }
        g[(int64_t)r * Cv] = acc;
    }
}

pf_batch toep_batch(const pf_toep* t, int C, const float2* b0s) {
    pf_batch b;
    b.C = C;
    b.Cv = 1;
    for (int v = std::min(64, C); v >= 1; --v)
        if (C % v == 0) {
            b.Cv = v;
            break;
        }
    b.Cb = C / b.Cv;
    b.L = t->L;
    b.K = t->K;
    b.R = t->R;
    b.nfg = t->nfr;
    b.nimg = t->nfr > 1 ? t->nfr : t->K;
    b.basis = nullptr;
    b.b0t = nullptr;
    b.b0s = b0s;
    b.nt = 1;
    return b;
}

size_t align_up(size_t n) { return (n + 255) / 256 * 256; }

size_t extra_bytes(const pf_toep* t, const pf_batch& b) {
    return pf_fftk_supported(t->geo) ? pf_fftk_extra_bytes(t->geo, b) : 0;
}

int toep_create(const void* kernel, int ndim, const int* shape, int nfr, int nseg, int ncoeff, float sigma,
                void* stream, pf_toep** out) {
    if (!out) return pf_fail(PF_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!kernel || !shape) return pf_fail(PF_ERR_ARG, "null argument");
    if (nseg < 1 || ncoeff < 1 || nfr < 1) return pf_fail(PF_ERR_ARG, "nframes, nseg and ncoeff must be >= 1");
    if (nfr > 1 && ncoeff > 1) return pf_fail(PF_ERR_ARG, "per-frame kernels and a subspace basis exclude each other");
    const int R = nseg * ncoeff;
    if (R > PF_TOEP_RMAX)
        return pf_fail(PF_ERR_UNSUPPORTED, "Toeplitz kernel blocks larger than " + std::to_string(PF_TOEP_RMAX) + "x" +
                                               std::to_string(PF_TOEP_RMAX) + " (nseg * ncoeff)");
    pf_toep* t = new (std::nothrow) pf_toep();
    if (!t) return pf_fail(PF_ERR_ALLOC, "host allocation failed");
    t->R = R;
    t->L = nseg;
    t->K = ncoeff;
    t->nfr = nfr;
    t->geo = new (std::nothrow) pf_plan();
    int w[3] = {1, 1, 1};
    float s[3] = {sigma, sigma, sigma};
    int rc = t->geo ? pf_plan_params(t->geo, ndim, shape, w, s, 1, 1, nullptr, 0) : pf_fail(PF_ERR_ALLOC, "host allocation");
    if (rc == PF_OK) rc = pf_plan_upload_apod(t->geo);
    if (rc == PF_OK) rc = pf_fftk_supported(t->geo) ? pf_fftk_init(t->geo) : pf_fft_cache_init(t->geo);
    if (rc == PF_OK) {
        const size_t n = (size_t)nfr * pf_gtot(t->geo) * R * R * sizeof(float2);
        hipError_t e = hipMalloc(&t->T, n);
        if (e == hipSuccess)
            e = hipMemcpyAsync(t->T, kernel, n, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream));
        if (e == hipSuccess) e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
        if (e != hipSuccess) rc = pf_fail(PF_ERR_HIP, std::string("Toeplitz kernel upload: ") + hipGetErrorString(e));
    }
    if (rc != PF_OK) {
        if (t->T) (void)hipFree(t->T);
        if (t->geo) pf_plan_free(t->geo);
        delete t;
        return rc;
    }
    *out = t;
    return PF_OK;
}

}  // namespace

extern "C" {

int pf_toeplitz_create(const void* kernel, int ndim, const int* shape, int nseg, int ncoeff, void* stream,
                       pf_toep** out) {
    return toep_create(kernel, ndim, shape, 1, nseg, ncoeff, 2.0f, stream, out);
}

int pf_gram_create(const void* kernel, int ndim, const int* shape, int nframes, int nseg, int ncoeff, void* stream,
                   pf_toep** out) {
    return toep_create(kernel, ndim, shape, nframes, nseg, ncoeff, 1.0f, stream, out);
}

int pf_toeplitz_destroy(pf_toep* t) {
    if (!t) return PF_OK;
    if (t->T) (void)hipFree(t->T);
    if (t->geo) pf_plan_free(t->geo);
    delete t;
    return PF_OK;
}

int pf_toeplitz_workspace_size(const pf_toep* t, int batch, size_t* bytes) {
    if (!t || !bytes) return pf_fail(PF_ERR_ARG, "null argument");
    if (batch < 1) return pf_fail(PF_ERR_ARG, "batch must be >= 1");
    pf_batch b = toep_batch(t, batch, nullptr);
    *bytes = align_up((size_t)b.nfg * b.C * b.R * pf_gtot(t->geo) * sizeof(float2)) + align_up(extra_bytes(t, b));
    return PF_OK;
}

int pf_toeplitz_apply(pf_toep* t, int batch, const void* image, void* out, const void* b0_spatial, void* workspace,
                      size_t workspace_bytes, void* stream) {
    if (!t || !image || !out) return pf_fail(PF_ERR_ARG, "null argument");
    if (batch < 1) return pf_fail(PF_ERR_ARG, "batch must be >= 1");
    if (t->L > 1 && !b0_spatial) return pf_fail(PF_ERR_ARG, "B0 segments need the spatial maps");
    pf_batch b = toep_batch(t, batch, static_cast<const float2*>(b0_spatial));
    const size_t gb = align_up((size_t)b.nfg * b.C * b.R * pf_gtot(t->geo) * sizeof(float2));
    const size_t need = gb + align_up(extra_bytes(t, b));
    if (!workspace || workspace_bytes < need) return pf_fail(PF_ERR_ARG, "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    float2* grid = static_cast<float2*>(workspace);
    float2* extra = reinterpret_cast<float2*>(static_cast<char*>(workspace) + gb);
    const bool fused = pf_fftk_supported(t->geo);
    int rc;
    if (fused) {
        rc = pf_fftk_forward(t->geo, b, static_cast<const float2*>(image), grid, extra, s);
    } else {
        // grid sizes with other prime factors: generic pad, rocFFT, crop
        rc = pf_launch_deapod_pad(t->geo, b, static_cast<const float2*>(image), grid, s);
        if (rc == PF_OK) rc = pf_fft_exec(t->geo, b, grid, 0, s);
    }
    if (rc != PF_OK) return rc;
    const int64_t gtot = pf_gtot(t->geo);
    const int64_t total = (int64_t)b.nfg * b.Cb * gtot * b.Cv;
    const unsigned nb = (unsigned)std::min<int64_t>((total + TB - 1) / TB, 1 << 20);
    if (t->R == 1)
        hipLaunchKernelGGL(k_toep_mul<1>, dim3(nb), dim3(TB), 0, s, t->T, grid, gtot, 1, b.Cv, b.Cb, b.nfg);
    else if (t->R <= 4)
        hipLaunchKernelGGL(k_toep_mul<4>, dim3(nb), dim3(TB), 0, s, t->T, grid, gtot, t->R, b.Cv, b.Cb, b.nfg);
    else if (t->R <= 16)
        hipLaunchKernelGGL(k_toep_mul<16>, dim3(nb), dim3(TB), 0, s, t->T, grid, gtot, t->R, b.Cv, b.Cb, b.nfg);
    else {
        const int64_t ngen = (total + TB_GEN - 1) / TB_GEN;
        if (ngen > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "grid too large for the generic Toeplitz multiply");
        hipLaunchKernelGGL(k_toep_mul_lds, dim3((unsigned)ngen), dim3(TB_GEN), (size_t)t->R * TB_GEN * sizeof(float2), s,
                           t->T, grid, gtot, t->R, b.Cv, b.Cb, b.nfg);
    }
    PF_HIP(hipGetLastError());
    const float factor = (float)(1.0 / (double)gtot);
    if (fused) return pf_fftk_adjoint(t->geo, b, grid, static_cast<float2*>(out), extra, factor, s);
    if ((rc = pf_fft_exec(t->geo, b, grid, 1, s)) != PF_OK) return rc;
    return pf_launch_crop_deapod(t->geo, b, grid, static_cast<float2*>(out), factor, s);
}

}  // extern "C"
