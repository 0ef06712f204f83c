// extern "C" entry points of libpfhip (declared in include/pfhip.h).
#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "../../include/pfhip.h"
#include "pf_internal.h"


static thread_local std::string g_err;

void pf_set_error(const std::string& msg) { g_err = msg; }
int pf_fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

namespace {

const size_t ALIGN = 256;
size_t align_up(size_t n) { return (n + ALIGN - 1) / ALIGN * ALIGN; }

// Grid layout of a call: the forward grid of batches the LDS-staged gather
// runs on is coil-pair major (pf_gather_col_layout); every other grid keeps
// all batch items innermost.  Entry points of the forward pipeline pass FWD,
// so separate stage calls (pad+FFT, interp, interp_grad) agree on it.
enum { DIR_ANY = 0, DIR_FWD = 1 };

int make_batch(const pf_plan* p, const pf_batch_desc* d, pf_batch* b, int dir = DIR_ANY) {
    if (!p) return pf_fail(PF_ERR_ARG, "null plan");
    if (!d) return pf_fail(PF_ERR_ARG, "null batch description");
    if (d->batch < 1) return pf_fail(PF_ERR_ARG, "batch must be >= 1");
    if (d->nseg < 1) return pf_fail(PF_ERR_ARG, "nseg must be >= 1");
    if (d->ncoeff < 0) return pf_fail(PF_ERR_ARG, "ncoeff must be >= 0");
    if (d->ncoeff > 0 && !d->basis) return pf_fail(PF_ERR_ARG, "ncoeff > 0 needs a basis");
    const bool has_b0 = d->b0_temporal || d->b0_spatial;
    if ((d->nseg > 1 || has_b0) && (!d->b0_temporal || !d->b0_spatial))
        return pf_fail(PF_ERR_ARG, "a field map needs both b0_temporal and b0_spatial");
    if (has_b0 && d->nt < 1) return pf_fail(PF_ERR_ARG, "nt must be >= 1");
    if (has_b0 && p->foff) return pf_fail(PF_ERR_UNSUPPORTED, "field-map segments need equal-length frames");
    b->C = d->batch;
    // all batch items innermost (one contiguous vector per grid point) up to
    // 64; larger batches split into blocks of the largest divisor <= 64
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
    if (b->R > 4096) return pf_fail(PF_ERR_UNSUPPORTED, "nseg*ncoeff too large");
    b->basis = d->ncoeff > 0 ? static_cast<const float2*>(d->basis) : nullptr;
    b->b0t = has_b0 ? static_cast<const float2*>(d->b0_temporal) : nullptr;
    b->b0s = has_b0 ? static_cast<const float2*>(d->b0_spatial) : nullptr;
    b->nt = has_b0 ? d->nt : 1;
    b->nfg = b->basis ? 1 : p->nframes;
    b->nimg = b->basis ? b->K : p->nframes;
    if (dir == DIR_FWD && pf_gather_col_layout(p, *b)) {
        b->Cv = 2;
        b->Cb = b->C / 2;
    }
    return PF_OK;
}

size_t grid_bytes(const pf_plan* p, const pf_batch& b) {
    return (size_t)b.nfg * b.C * b.R * (size_t)pf_gtot(p) * sizeof(float2);
}
size_t samples_bytes(const pf_plan* p, const pf_batch& b) { return (size_t)p->mtot * b.C * sizeof(float2); }

double ghat(const pf_plan* p) { return (double)pf_gtot(p); }

// workspace of the fused operators: [grid | sorted samples | FFT scratch | spread partial tiles]
struct WsLayout {
    size_t grid, samples, fft, spread, fwd_total, adj_total;
};

// The adjoint of a B0 / basis batch whose R*C values form one vector sorts
// them expanded (pf_spread_expand): its sorted samples are (Mtot, R*C) and its
// spread workspace that of the plain R*C-item batch.  Forward and adjoint
// share one layout.
WsLayout layout(const pf_plan* p, const pf_batch& b) {
    WsLayout w;
    pf_batch bx;
    const bool ex = pf_spread_expand(p, b, &bx);
    w.grid = 0;
    w.samples = align_up(grid_bytes(p, b));
    w.fft = w.samples + align_up(ex ? std::max(samples_bytes(p, b), samples_bytes(p, bx)) : samples_bytes(p, b));
    size_t extra = pf_fftk_supported(p) ? pf_fftk_extra_bytes(p, b) : 0;
    w.spread = w.fft + align_up(extra);
    w.fwd_total = w.spread;
    w.adj_total = w.spread + align_up(pf_spread_partial_bytes(p, ex ? bx : b));
    return w;
}

// FFT state a plan needs before its first operator call: the fused pruned
// FFT's twiddle tables for grid sizes 2^a 3^b 5^c (built here, so calls
// neither allocate nor synchronise), otherwise the rocFFT plan cache, whose
// per-batch-size rocFFT plans are made on first use under its lock.
int plan_fft_state(pf_plan* p) {
    int rc = pf_ops_prepare();
    if (rc != PF_OK) return rc;
    // the no-allocation entry point's scratch address on this device; a
    // failure here only disables that entry point (it then reports it)
    if (pf_noalloc_prepare() != PF_OK) pf_set_error("");
    if (pf_fftk_supported(p)) return pf_fftk_init(p);
    return pf_fft_cache_init(p);
}

}  // namespace

// Deapodisation vectors per axis (src/pyfourier/_subroutines/_apod.py:27-41):
// a(i) = sqrt(beta^2 - (pi W (i - n//2) / G)^2), a/sinh(a), normalised at
// n//2, NaN -> 1; all ones on width-1 axes.
int pf_plan_upload_apod(pf_plan* p) {
    int ntot = p->shape[0] + p->shape[1] + p->shape[2];
    std::vector<float> h(ntot, 1.0f);
    int o = 0;
    for (int a = 0; a < 3; ++a) {
        int n = p->shape[a];
        if (p->width[a] > 1) {
            auto raw = [&](int i) {
                double t = M_PI * p->width[a] * (double)(i - n / 2) / p->os[a];
                double v = std::sqrt((double)p->beta[a] * p->beta[a] - t * t);
                return v / std::sinh(v);
            };
            double c = raw(n / 2);
            for (int i = 0; i < n; ++i) {
                double v = raw(i) / c;
                h[o + i] = std::isnan(v) ? 1.0f : (float)v;
            }
        }
        o += n;
    }
    hipError_t he = hipMalloc(&p->apod, ntot * sizeof(float));
    if (he == hipSuccess) he = hipMemcpy(p->apod, h.data(), ntot * sizeof(float), hipMemcpyHostToDevice);
    if (he != hipSuccess) return pf_fail(PF_ERR_HIP, std::string("apod upload: ") + hipGetErrorString(he));
    return PF_OK;
}

extern "C" {

const char* pf_last_error(void) { return g_err.c_str(); }

int pf_version(void) { return 100; }

int pf_plan_create(const float* coord, int ndim, const int* shape, const int* width, const float* oversamp,
                   int nframes, long long npts, void* stream, pf_plan** out) {
    return pf_plan_create_tiled(coord, ndim, shape, width, oversamp, nframes, npts, nullptr, 0, stream, out);
}

int pf_plan_create_tiled(const float* coord, int ndim, const int* shape, const int* width, const float* oversamp,
                         int nframes, long long npts, const int* tile, int chunk, void* stream, pf_plan** out) {
    if (!out) return pf_fail(PF_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!coord || !shape || !width || !oversamp) return pf_fail(PF_ERR_ARG, "null argument");
    pf_plan* p = new (std::nothrow) pf_plan();
    if (!p) return pf_fail(PF_ERR_ALLOC, "host allocation failed");
    int rc = pf_plan_params(p, ndim, shape, width, oversamp, nframes, (int64_t)npts, tile, chunk);
    if (rc != PF_OK) {
        pf_plan_free(p);
        return rc;
    }
    if (hipGetDevice(&p->device) != hipSuccess) {
        pf_plan_free(p);
        return pf_fail(PF_ERR_HIP, "hipGetDevice failed");
    }
    if ((rc = pf_plan_upload_apod(p)) != PF_OK) {
        pf_plan_free(p);
        return rc;
    }
    rc = pf_plan_build(p, coord, static_cast<hipStream_t>(stream));
    if (rc == PF_OK) rc = plan_fft_state(p);
    if (rc != PF_OK) {
        pf_plan_free(p);
        return rc;
    }
    *out = p;
    return PF_OK;
}

int pf_plan_create_ragged(const float* coord, int ndim, const int* shape, const int* width, const float* oversamp,
                          int nframes, const long long* frame_offsets, const int* tile, int chunk, void* stream,
                          pf_plan** out) {
    if (!out) return pf_fail(PF_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!coord || !shape || !width || !oversamp || !frame_offsets) return pf_fail(PF_ERR_ARG, "null argument");
    if (nframes < 1) return pf_fail(PF_ERR_ARG, "nframes must be >= 1");
    if (frame_offsets[0] != 0) return pf_fail(PF_ERR_ARG, "frame_offsets[0] must be 0");
    for (int f = 0; f < nframes; ++f)
        if (frame_offsets[f + 1] < frame_offsets[f]) return pf_fail(PF_ERR_ARG, "frame_offsets must not decrease");
    const long long total = frame_offsets[nframes];
    pf_plan* p = new (std::nothrow) pf_plan();
    if (!p) return pf_fail(PF_ERR_ALLOC, "host allocation failed");
    int rc = pf_plan_params(p, ndim, shape, width, oversamp, 1, (int64_t)total, tile, chunk);
    if (rc == PF_OK && (int64_t)nframes * p->ntile_tot >= (1ll << 31))
        rc = pf_fail(PF_ERR_UNSUPPORTED, "too many tiles");
    if (rc != PF_OK) {
        pf_plan_free(p);
        return rc;
    }
    p->nframes = nframes;
    if (hipGetDevice(&p->device) != hipSuccess) {
        pf_plan_free(p);
        return pf_fail(PF_ERR_HIP, "hipGetDevice failed");
    }
    std::vector<int64_t> off(frame_offsets, frame_offsets + nframes + 1);
    hipError_t he = hipMalloc(&p->foff, off.size() * sizeof(int64_t));
    if (he == hipSuccess) he = hipMemcpy(p->foff, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        pf_plan_free(p);
        return pf_fail(PF_ERR_HIP, std::string("frame offsets: ") + hipGetErrorString(he));
    }
    if ((rc = pf_plan_upload_apod(p)) != PF_OK || (rc = pf_plan_build(p, coord, static_cast<hipStream_t>(stream))) != PF_OK ||
        (rc = plan_fft_state(p)) != PF_OK) {
        pf_plan_free(p);
        return rc;
    }
    *out = p;
    return PF_OK;
}

int pf_spread_tiles(int ndim, int* tiles, int cap, int* count) {
    if (ndim < 1 || ndim > 3) return pf_fail(PF_ERR_ARG, "ndim must be 1, 2 or 3");
    if (!count || cap < 0 || (cap > 0 && !tiles)) return pf_fail(PF_ERR_ARG, "null argument");
    *count = pf_spread_reg_tiles(ndim, tiles, cap);
    return PF_OK;
}

int pf_plan_destroy(pf_plan* plan) {
    pf_plan_free(plan);
    return PF_OK;
}

int pf_plan_get_info(const pf_plan* p, pf_plan_info* info) {
    if (!p || !info) return pf_fail(PF_ERR_ARG, "null argument");
    info->ndim = p->ndim;
    for (int a = 0; a < 3; ++a) {
        info->shape[a] = p->shape[a];
        info->os_shape[a] = p->os[a];
        info->width[a] = p->width[a];
        info->beta[a] = p->beta[a];
        info->tile[a] = p->tile[a];
    }
    info->scale = p->scale;
    info->nframes = p->nframes;
    info->npts = p->npts;
    info->spread_entries = p->sl[0].nslist;
    info->spread_items = p->sl[0].split[1].nsitems;
    info->spread_slots = p->sl[0].split[1].nslots;
    return PF_OK;
}

int pf_plan_get_permutation(const pf_plan* p, int* perm, void* stream) {
    if (!p || !perm) return pf_fail(PF_ERR_ARG, "null argument");
    PF_HIP(hipMemcpyAsync(perm, p->perm, p->mtot * sizeof(int), hipMemcpyDeviceToDevice,
                          static_cast<hipStream_t>(stream)));
    return PF_OK;
}

int pf_describe_launch(const pf_plan* p, const pf_batch_desc* d, char* buf, size_t cap) {
    pf_batch bf, ba;
    int rc = make_batch(p, d, &bf, DIR_FWD);
    if (rc == PF_OK) rc = make_batch(p, d, &ba);
    if (rc != PF_OK) return rc;
    if (!buf || cap == 0) return pf_fail(PF_ERR_ARG, "null output buffer");
    pf_batch bx;
    const std::string spread =
        pf_spread_expand(p, ba, &bx) ? "k_sort_x+" + pf_spread_name(p, bx) : pf_spread_name(p, ba);
    const std::string text = "gather=" + pf_gather_name(p, bf) + " spread=" + spread +
                             " fft=" + (pf_fftk_supported(p) ? "fused" : "rocfft");
    const size_t n = std::min(cap - 1, text.size());
    std::copy(text.begin(), text.begin() + n, buf);
    buf[n] = 0;
    return text.size() < cap ? PF_OK : pf_fail(PF_ERR_ARG, "output buffer too small");
}

int pf_grid_size(const pf_plan* p, const pf_batch_desc* d, size_t* bytes) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!bytes) return pf_fail(PF_ERR_ARG, "null output");
    *bytes = grid_bytes(p, b);
    return PF_OK;
}

int pf_spread_workspace_size(const pf_plan* p, const pf_batch_desc* d, size_t* bytes) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!bytes) return pf_fail(PF_ERR_ARG, "null output");
    *bytes = pf_spread_partial_bytes(p, b);
    return PF_OK;
}

int pf_workspace_size(const pf_plan* p, const pf_batch_desc* d, int op, size_t* bytes) {
    pf_batch b;
    int rc = make_batch(p, d, &b, op == 0 ? DIR_FWD : DIR_ANY);
    if (rc != PF_OK) return rc;
    if (!bytes) return pf_fail(PF_ERR_ARG, "null output");
    if (op != 0 && op != 1) return pf_fail(PF_ERR_ARG, "op must be 0 (forward) or 1 (adjoint)");
    WsLayout w = layout(p, b);
    *bytes = op == 0 ? w.fwd_total : w.adj_total;
    return PF_OK;
}

int pf_deapod_pad(const pf_plan* p, const pf_batch_desc* d, const void* image, void* grid, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b, DIR_FWD);
    if (rc != PF_OK) return rc;
    if (!image || !grid) return pf_fail(PF_ERR_ARG, "null array");
    return pf_launch_deapod_pad(p, b, static_cast<const float2*>(image), static_cast<float2*>(grid),
                                static_cast<hipStream_t>(stream));
}

int pf_fft(pf_plan* p, const pf_batch_desc* d, void* grid, int inverse, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b, inverse ? DIR_ANY : DIR_FWD);
    if (rc != PF_OK) return rc;
    if (!grid) return pf_fail(PF_ERR_ARG, "null array");
    return pf_fft_exec(p, b, static_cast<float2*>(grid), inverse ? 1 : 0, static_cast<hipStream_t>(stream));
}

int pf_interp(const pf_plan* p, const pf_batch_desc* d, const void* grid, void* ksp_sorted, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b, DIR_FWD);
    if (rc != PF_OK) return rc;
    if (!grid || !ksp_sorted) return pf_fail(PF_ERR_ARG, "null array");
    return pf_launch_gather(p, b, static_cast<const float2*>(grid), static_cast<float2*>(ksp_sorted),
                            static_cast<hipStream_t>(stream));
}

int pf_spread(const pf_plan* p, const pf_batch_desc* d, const void* ksp_sorted, void* grid, void* workspace,
              void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!grid || !ksp_sorted) return pf_fail(PF_ERR_ARG, "null array");
    return pf_launch_spread(p, b, static_cast<const float2*>(ksp_sorted), static_cast<float2*>(grid),
                            static_cast<float2*>(workspace), static_cast<hipStream_t>(stream));
}

int pf_crop_deapod(const pf_plan* p, const pf_batch_desc* d, const void* grid, void* image, float factor,
                   void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!image || !grid) return pf_fail(PF_ERR_ARG, "null array");
    return pf_launch_crop_deapod(p, b, static_cast<const float2*>(grid), static_cast<float2*>(image), factor,
                                 static_cast<hipStream_t>(stream));
}

int pf_sort_samples(const pf_plan* p, const pf_batch_desc* d, const void* kspace, void* ksp_sorted,
                    const float* weight, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!kspace || !ksp_sorted) return pf_fail(PF_ERR_ARG, "null array");
    return pf_launch_sort(p, b, static_cast<const float2*>(kspace), static_cast<float2*>(ksp_sorted), weight,
                          static_cast<hipStream_t>(stream));
}

int pf_unsort_samples(const pf_plan* p, const pf_batch_desc* d, const void* ksp_sorted, void* kspace,
                      const float* weight, float factor, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!kspace || !ksp_sorted) return pf_fail(PF_ERR_ARG, "null array");
    return pf_launch_unsort(p, b, static_cast<const float2*>(ksp_sorted), static_cast<float2*>(kspace), weight,
                            factor, static_cast<hipStream_t>(stream));
}

int pf_workspace_offsets(const pf_plan* p, const pf_batch_desc* d, size_t* grid_off, size_t* samples_off,
                         size_t* fft_off, size_t* spread_off) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!grid_off || !samples_off || !fft_off || !spread_off) return pf_fail(PF_ERR_ARG, "null output");
    WsLayout w = layout(p, b);
    *grid_off = w.grid;
    *samples_off = w.samples;
    *fft_off = w.fft;
    *spread_off = w.spread;
    return PF_OK;
}

int pf_interp_grad(const pf_plan* p, const pf_batch_desc* d, const void* grid, const void* h_sorted,
                   const float* axis_scale, int accumulate, float* dcoord, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b, DIR_FWD);
    if (rc != PF_OK) return rc;
    if (!grid || !h_sorted || !dcoord) return pf_fail(PF_ERR_ARG, "null array");
    float sc[3] = {1.f, 1.f, 1.f};
    if (axis_scale)
        for (int a = 0; a < 3; ++a) sc[a] = axis_scale[a];
    return pf_launch_gather_dcoord(p, b, static_cast<const float2*>(grid), static_cast<const float2*>(h_sorted), sc,
                                   accumulate != 0, dcoord, static_cast<hipStream_t>(stream));
}

int pf_pad_fft(pf_plan* p, const pf_batch_desc* d, const void* image, void* grid, void* fft_workspace,
               void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b, DIR_FWD);
    if (rc != PF_OK) return rc;
    if (!image || !grid) return pf_fail(PF_ERR_ARG, "null array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (pf_fftk_supported(p)) {
        if (pf_fftk_extra_bytes(p, b) && !fft_workspace) return pf_fail(PF_ERR_ARG, "missing FFT workspace");
        return pf_fftk_forward(p, b, static_cast<const float2*>(image), static_cast<float2*>(grid),
                               static_cast<float2*>(fft_workspace), s);
    }
    if ((rc = pf_launch_deapod_pad(p, b, static_cast<const float2*>(image), static_cast<float2*>(grid), s)) != PF_OK)
        return rc;
    return pf_fft_exec(p, b, static_cast<float2*>(grid), 0, s);
}

int pf_ifft_crop(pf_plan* p, const pf_batch_desc* d, void* grid, void* image, float factor, void* fft_workspace,
                 void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!image || !grid) return pf_fail(PF_ERR_ARG, "null array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (pf_fftk_supported(p)) {
        if (pf_fftk_extra_bytes(p, b) && !fft_workspace) return pf_fail(PF_ERR_ARG, "missing FFT workspace");
        return pf_fftk_adjoint(p, b, static_cast<float2*>(grid), static_cast<float2*>(image),
                               static_cast<float2*>(fft_workspace), factor, s);
    }
    if ((rc = pf_fft_exec(p, b, static_cast<float2*>(grid), 1, s)) != PF_OK) return rc;
    return pf_launch_crop_deapod(p, b, static_cast<const float2*>(grid), static_cast<float2*>(image), factor, s);
}

int pf_nufft(pf_plan* p, const pf_batch_desc* d, const void* image, void* kspace, const float* weight, int norm,
             void* workspace, size_t workspace_bytes, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b, DIR_FWD);
    if (rc != PF_OK) return rc;
    if (!image || !kspace) return pf_fail(PF_ERR_ARG, "null array");
    if (norm < PF_NORM_NONE || norm > PF_NORM_ADJOINT) return pf_fail(PF_ERR_ARG, "norm must be 0, 1 or 2");
    WsLayout w = layout(p, b);
    if (!workspace || workspace_bytes < w.fwd_total) return pf_fail(PF_ERR_ARG, "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace);
    float2* grid = reinterpret_cast<float2*>(ws + w.grid);
    float2* ys = reinterpret_cast<float2*>(ws + w.samples);
    float2* extra = reinterpret_cast<float2*>(ws + w.fft);
    double factor = 1.0 / p->scale;
    if (norm == PF_NORM_ORTHO) factor /= std::sqrt(ghat(p));
    if (norm == PF_NORM_ADJOINT) factor /= ghat(p);
    if (pf_fftk_supported(p)) {
        if ((rc = pf_fftk_forward(p, b, static_cast<const float2*>(image), grid, extra, s)) != PF_OK) return rc;
    } else {
        if ((rc = pf_launch_deapod_pad(p, b, static_cast<const float2*>(image), grid, s)) != PF_OK) return rc;
        if ((rc = pf_fft_exec(p, b, grid, 0, s)) != PF_OK) return rc;
    }
    if ((rc = pf_launch_gather(p, b, grid, ys, s)) != PF_OK) return rc;
    return pf_launch_unsort(p, b, ys, static_cast<float2*>(kspace), weight, (float)factor, s);
}

int pf_nufft_adj(pf_plan* p, const pf_batch_desc* d, const void* kspace, void* image, const float* weight,
                 int norm, void* workspace, size_t workspace_bytes, void* stream) {
    pf_batch b;
    int rc = make_batch(p, d, &b);
    if (rc != PF_OK) return rc;
    if (!image || !kspace) return pf_fail(PF_ERR_ARG, "null array");
    if (norm < PF_NORM_NONE || norm > PF_NORM_ADJOINT) return pf_fail(PF_ERR_ARG, "norm must be 0, 1 or 2");
    WsLayout w = layout(p, b);
    if (!workspace || workspace_bytes < w.adj_total) return pf_fail(PF_ERR_ARG, "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace);
    float2* grid = reinterpret_cast<float2*>(ws + w.grid);
    float2* ys = reinterpret_cast<float2*>(ws + w.samples);
    float2* extra = reinterpret_cast<float2*>(ws + w.fft);
    float2* part = reinterpret_cast<float2*>(ws + w.spread);
    double factor = 1.0 / p->scale;
    if (norm == PF_NORM_ORTHO) factor /= std::sqrt(ghat(p));
    if (norm == PF_NORM_NONE) factor /= ghat(p);
    // 3-D plans with a column mask: the spread leaves the empty tiles of
    // columns no sample covers unwritten, and the inverse FFT takes those
    // columns as zero without reading them
    const bool masked = pf_fftk_supported(p) && pf_fftk_masked(p);
    pf_batch bx;
    if (pf_spread_expand(p, b, &bx)) {
        // B0 segments / basis coefficients applied in the sort: a plain spread of R*C items
        if ((rc = pf_launch_sort_expand(p, b, static_cast<const float2*>(kspace), ys, weight, s)) != PF_OK) return rc;
        if ((rc = pf_launch_spread(p, bx, ys, grid, part, s, false, masked)) != PF_OK) return rc;
    } else {
        // the sort writes the spread's per-sample maxima when the spread uses them
        bool smax_ready = false;
        if ((rc = pf_launch_sort(p, b, static_cast<const float2*>(kspace), ys, weight, s, pf_spread_smax(p, b, part),
                                 &smax_ready)) != PF_OK)
            return rc;
        if ((rc = pf_launch_spread(p, b, ys, grid, part, s, smax_ready, masked)) != PF_OK) return rc;
    }
    if (pf_fftk_supported(p))
        return pf_fftk_adjoint(p, b, grid, static_cast<float2*>(image), extra, (float)factor, s, masked);
    if ((rc = pf_fft_exec(p, b, grid, 1, s)) != PF_OK) return rc;
    return pf_launch_crop_deapod(p, b, grid, static_cast<float2*>(image), (float)factor, s);
}

}  // extern "C"
