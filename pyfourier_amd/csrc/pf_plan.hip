// Plan construction: per-sample Kaiser-Bessel tap records, tile binning and a
// stable sort of the trajectory, and the spread work lists.
//
// Follows the reference interpolator plan
//   src/pyfourier/_nufft/_plan.py:117-155            (beta, oversampled shape)
//   src/pyfourier/_subroutines/_traj.py:57-66         (c_os = c*G/N + G//2, float32)
//   src/pyfourier/_subroutines/_grid/_plan.py:157-210 (taps, KB values, mod G)
//   src/pyfourier/_subroutines/_grid/_plan.py:118-141 (kernel scale)
// but stores one compact record per sample (first stored tap + W weights per
// axis) instead of (index, value) tables, in tile-sorted order.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>


#include "pf_internal.h"

namespace {

// Kaiser-Bessel I0(beta*sqrt(1-x^2)) with the Abramowitz & Stegun 9.8.1 and
// 9.8.2 polynomial approximations (the ones the reference evaluates), 0 for
// |x| > 1.  Double precision; the tap value is rounded to float afterwards,
// as the reference stores float32 tables.
__host__ __device__ double kb_value(double x, double beta) {
    if (fabs(x) > 1.0) return 0.0;
    double u = beta * sqrt(1.0 - x * x);
    double t = u / 3.75;
    if (u < 3.75) {
        double t2 = t * t;
        return 1.0 + t2 * (3.5156229 + t2 * (3.0899424 + t2 * (1.2067492 +
               t2 * (0.2659732 + t2 * (0.0360768 + t2 * 0.0045813)))));
    }
    double it = 1.0 / t;
    double p = 0.39894228 + it * (0.01328592 + it * (0.00225319 + it * (-0.00157565 +
               it * (0.00916281 + it * (-0.02057706 + it * (0.02635537 +
               it * (-0.01647633 + it * 0.00392377)))))));
    return p * exp(u) / sqrt(u);
}

struct TapParams {
    int ndim;
    int os[3];
    int width[3];
    float beta[3];
    float sc[3];   // G/N in float32
    float sh[3];   // G//2 in float32
    int tile[3];
    int ntiles[3];
    int64_t npts;
    int64_t mtot;
    int64_t ntile_tot;
    int rec;
    int woff[3];
    int mbits[3];  // Morton bits per axis (0 on all axes: row-major tile order)
    int ibits[3];  // Morton bits of the position inside the tile (0: tile order only)
    const int64_t* foff;  // ragged frame offsets (nframes + 1) or null
    int nframes;
};

// Interleave the bits of (tz, ty, tx), x fastest: consecutive codes are
// spatially compact blocks of tiles (Z-order), which keeps the halo of the
// tiles in flight within L2.
__host__ __device__ inline uint32_t morton3(uint32_t tz, uint32_t ty, uint32_t tx, const int* mb) {
    uint32_t code = 0;
    int o = 0;
    const int mx = mb[0] > mb[1] ? (mb[0] > mb[2] ? mb[0] : mb[2]) : (mb[1] > mb[2] ? mb[1] : mb[2]);
    for (int b = 0; b < mx; ++b) {
        if (b < mb[2]) code |= ((tx >> b) & 1u) << o++;
        if (b < mb[1]) code |= ((ty >> b) & 1u) << o++;
        if (b < mb[0]) code |= ((tz >> b) & 1u) << o++;
    }
    return code;
}

__device__ __forceinline__ int64_t pmod(int64_t a, int64_t m) {
    int64_t r = a % m;
    return r < 0 ? r + m : r;
}

// One thread per original sample: tap record and tile key.
__global__ void k_taps(const float* __restrict__ coord, TapParams P, uint32_t* __restrict__ keys,
                       int32_t* __restrict__ ids, float* __restrict__ rec,
                       float* __restrict__ cos_out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.mtot) return;
    int64_t f;
    if (P.foff) {
        // last frame whose first sample is <= i
        int lo = 0, hi = P.nframes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (P.foff[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        f = lo;
    } else {
        f = i / P.npts;
    }
    int64_t tcoord[3];
    float* r = rec + i * P.rec;
    int* ri = reinterpret_cast<int*>(r);
    for (int a = 0; a < 3; ++a) {
        int W = P.width[a];
        int G = P.os[a];
        int64_t x0;
        if (a < 3 - P.ndim) {
            x0 = 0;
            r[P.woff[a]] = 1.0f;
            cos_out[i * 3 + a] = 0.0f;
            ri[a] = 0;
            tcoord[a] = 0;
            continue;
        }
        // coordinate axis order is (x, y, z): grid slot 2 <- coord[..., 0]
        float c = coord[i * P.ndim + (2 - a)];
        // two roundings like the reference's `c *= scale; c += shift`
        // (_subroutines/_traj.py:64-65); a fused multiply-add would round once
        // and move samples that sit on a tap boundary to the next tap
        float prod;
        {
#pragma clang fp contract(off)
            prod = c * P.sc[a];
        }
        float cos_;
        {
#pragma clang fp contract(off)
            cos_ = prod + P.sh[a];
        }
        cos_out[i * 3 + a] = cos_;
        double cd = (double)cos_;
        if (W == 1) {
            x0 = (int64_t)rint(cd);
            r[P.woff[a]] = 1.0f;
        } else {
            double half = 0.5 * (double)W;
            x0 = (int64_t)ceil(cd - half);
            for (int k = 0; k < W; ++k)
                r[P.woff[a] + k] = (float)kb_value(((double)(x0 + k) - cd) / half, (double)P.beta[a]);
        }
        // centred index j is stored at (j - G//2) mod G
        int64_t st = pmod(x0 - G / 2, G);
        ri[a] = (int)st;
        tcoord[a] = st / P.tile[a];
    }
    ri[3] = (int)f;
    const int msum = P.mbits[0] + P.mbits[1] + P.mbits[2];
    const int isum = P.ibits[0] + P.ibits[1] + P.ibits[2];
    if (msum > 0)
        // Z order of the tiles, then Z order inside the tile: consecutive
        // samples have overlapping footprints (gather cache reuse)
        keys[i] = ((uint32_t)f << (msum + isum)) |
                  (morton3((uint32_t)tcoord[0], (uint32_t)tcoord[1], (uint32_t)tcoord[2], P.mbits) << isum) |
                  morton3((uint32_t)(ri[0] - tcoord[0] * P.tile[0]), (uint32_t)(ri[1] - tcoord[1] * P.tile[1]),
                          (uint32_t)(ri[2] - tcoord[2] * P.tile[2]), P.ibits);
    else
        keys[i] = (uint32_t)(f * P.ntile_tot +
                             (tcoord[0] * P.ntiles[1] + tcoord[1]) * P.ntiles[2] + tcoord[2]);
    ids[i] = (int32_t)i;
}

// Gather records into sorted order and build the inverse permutation.
__global__ void k_permute(const int32_t* __restrict__ perm, const float* __restrict__ rec_in,
                          float* __restrict__ rec_out, const float* __restrict__ cos_in,
                          float* __restrict__ cos_out, int32_t* __restrict__ iperm, int64_t mtot,
                          int rec) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= mtot) return;
    int64_t i = perm[s];
    iperm[i] = (int32_t)s;
    const int* ri = reinterpret_cast<const int*>(rec_in) + i * rec;
    int* ro = reinterpret_cast<int*>(rec_out) + s * rec;
    for (int k = 0; k < rec; ++k) ro[k] = ri[k];
    for (int k = 0; k < 3; ++k) cos_out[s * 3 + k] = cos_in[i * 3 + k];
}

// Distinct tiles touched along one axis by taps start..start+W-1 (mod G).
__device__ __forceinline__ int axis_tiles(int start, int W, int G, int T, int* out) {
    int n = 0;
    int prev = -1;
    for (int k = 0; k < W; ++k) {
        int g = start + k;
        if (g >= G) g -= G;
        int t = g / T;
        if (t != prev) {
            bool seen = false;
            for (int q = 0; q < n; ++q) seen |= (out[q] == t);
            if (!seen) out[n++] = t;
            prev = t;
        }
    }
    return n;
}

__global__ void k_slist_count(const float* __restrict__ rec, TapParams P, uint32_t* __restrict__ cnt) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.mtot) return;
    const int* ri = reinterpret_cast<const int*>(rec + s * P.rec);
    int t[PF_WMAX];
    uint32_t n = 1;
    for (int a = 0; a < 3; ++a) n *= (uint32_t)axis_tiles(ri[a], P.width[a], P.os[a], P.tile[a], t);
    cnt[s] = n;
}

__global__ void k_slist_emit(const float* __restrict__ rec, TapParams P, int merge, const uint32_t* __restrict__ off,
                             uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.mtot) return;
    const int* ri = reinterpret_cast<const int*>(rec + s * P.rec);
    int tz[PF_WMAX], ty[PF_WMAX], tx[PF_WMAX];
    int nz = axis_tiles(ri[0], P.width[0], P.os[0], P.tile[0], tz);
    int ny = axis_tiles(ri[1], P.width[1], P.os[1], P.tile[1], ty);
    int nx = axis_tiles(ri[2], P.width[2], P.os[2], P.tile[2], tx);
    int64_t f = merge ? 0 : ri[3];
    uint32_t o = off[s];
    for (int a = 0; a < nz; ++a)
        for (int b = 0; b < ny; ++b)
            for (int c = 0; c < nx; ++c) {
                keys[o] = (uint32_t)(f * P.ntile_tot + ((int64_t)tz[a] * P.ntiles[1] + ty[b]) * P.ntiles[2] + tx[c]);
                vals[o] = (int32_t)s;
                ++o;
            }
}

__global__ void k_bounds(const uint32_t* __restrict__ keys, int64_t n, int32_t* __restrict__ start,
                         int32_t* __restrict__ end) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    uint32_t k = keys[e];
    if (e == 0 || keys[e - 1] != k) start[k] = (int32_t)e;
    if (e == n - 1 || keys[e + 1] != k) end[k] = (int32_t)(e + 1);
}

// Window key of every sorted sample for the LDS-staged gather: (frame,
// column of its first y tap, column of its first x tap, window of its first z
// tap), and the LDS index of its first tap inside that window (the window is
// (DZ + W - 1) planes of (PF_GCOL + W - 1)^2 points); the values are the
// sorted indices themselves.
__global__ void k_gather_keys(const float* __restrict__ rec, TapParams P, int ncy, int ncx, int nzw, int dz,
                              uint32_t* __restrict__ keys, int32_t* __restrict__ vals, int32_t* __restrict__ base) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.mtot) return;
    const int* ri = reinterpret_cast<const int*>(rec + t * P.rec);
    const int64_t f = ri[3];
    keys[t] = (uint32_t)((((f * ncy) + ri[1] / PF_GCOL) * ncx + ri[2] / PF_GCOL) * nzw + ri[0] / dz);
    vals[t] = (int32_t)t;
    const int px = PF_GCOL + P.width[2] - 1;
    base[t] = ((ri[0] % dz) * px + ri[1] % PF_GCOL) * px + ri[2] % PF_GCOL;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 1); }
    template <class T>
    T* as() {
        return static_cast<T*>(p);
    }
};

unsigned bits_for(uint64_t n) {
    unsigned b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

double pf_kb_value_host(double x, double beta) { return kb_value(x, beta); }

// Fill the scalar plan parameters; coordinates are in grid units.
int pf_plan_params(pf_plan* p, int ndim, const int* shape_zyx, const int* width_zyx,
                   const float* oversamp_zyx, int nframes, int64_t npts, const int* tile_zyx, int chunk) {
    if (ndim < 1 || ndim > 3) return pf_fail(PF_ERR_ARG, "ndim must be 1, 2 or 3");
    if (nframes < 1 || npts < 1) return pf_fail(PF_ERR_ARG, "empty trajectory");
    if ((int64_t)nframes * npts >= (1ll << 31))
        return pf_fail(PF_ERR_UNSUPPORTED, "more than 2^31-1 samples in one plan");
    p->ndim = ndim;
    p->nframes = nframes;
    p->npts = npts;
    p->mtot = (int64_t)nframes * npts;
    double scale = 1.0;
    for (int a = 0; a < 3; ++a) {
        int src = a - (3 - ndim);
        if (src < 0) {
            p->shape[a] = 1;
            p->os[a] = 1;
            p->width[a] = 1;
            p->oversamp[a] = 1.0f;
            p->beta[a] = 0.0f;
            continue;
        }
        int N = shape_zyx[src];
        int W = width_zyx[src];
        float s = oversamp_zyx[src];
        if (N < 1) return pf_fail(PF_ERR_ARG, "image size must be positive");
        if (W < 1 || W > PF_WMAX) return pf_fail(PF_ERR_UNSUPPORTED, "kernel width must be in [1, 8]");
        if (!(s >= 1.0f)) return pf_fail(PF_ERR_ARG, "oversampling must be >= 1");
        p->shape[a] = N;
        p->width[a] = W;
        p->oversamp[a] = s;
        // beta = pi*sqrt((W/s*(s-0.5))^2 - 0.8) in float32 (_nufft/_plan.py:128)
        float q = ((float)W / s) * (s - 0.5f);
        float inner = q * q - 0.8f;
        p->beta[a] = (float)M_PI * sqrtf(inner);
        p->os[a] = (int)ceilf(s * (float)N);
        if (p->os[a] < W) return pf_fail(PF_ERR_ARG, "oversampled grid smaller than the kernel");
        if (W > 1) {
            // sum of KB at the half-open centred taps (_grid/_plan.py:118-141)
            double start = std::ceil(-W / 2.0), sum = 0.0;
            for (int e = 0; e < W; ++e) sum += kb_value((start + e) / (W / 2.0), (double)p->beta[a]);
            scale *= sum;
        }
    }
    p->scale = scale;
    // spread/sort tiles: the register-resident spread holds one tile for
    // 32 batch values in one wave (k_spread_reg)
    // 3-D default 8x4x4: narrow in x, so a lane's two x points are mostly
    // inside a sample's footprint (measured fastest on radial 3-D, W 4 and 6)
    int t3[3] = {8, 4, 4}, t2[3] = {1, 8, 16}, t1[3] = {1, 1, 64};
    int tu[3] = {1, 1, 1};
    const int* tt = ndim == 3 ? t3 : (ndim == 2 ? t2 : t1);
    if (tile_zyx) {
        // caller-chosen tile (ndim entries, z, y, x order), e.g. by an autotuner
        for (int a = 3 - ndim; a < 3; ++a) {
            const int t = tile_zyx[a - (3 - ndim)];
            if (t < 1 || t > 4096) return pf_fail(PF_ERR_ARG, "tile extents must be in [1, 4096]");
            tu[a] = t;
        }
        tt = tu;
    }
    p->ntile_tot = 1;
    int off = 4;
    for (int a = 0; a < 3; ++a) {
        p->tile[a] = std::min(tt[a], p->os[a]);
        // a footprint must not leave a tile and re-enter it across the
        // periodic wrap: with several tiles that needs G >= T + W
        if (p->os[a] < p->tile[a] + p->width[a]) p->tile[a] = p->os[a];
        p->ntiles[a] = (p->os[a] + p->tile[a] - 1) / p->tile[a];
        p->ntile_tot *= p->ntiles[a];
        p->woff[a] = off;
        off += p->width[a];
    }
    if ((int64_t)nframes * p->ntile_tot >= (1ll << 31))
        return pf_fail(PF_ERR_UNSUPPORTED, "too many tiles");
    p->rec = (off + 3) & ~3;
    if (chunk < 0) return pf_fail(PF_ERR_ARG, "chunk must be >= 0 (0: automatic)");
    p->max_chunk = chunk;  // 0: chosen from the spread list size when it is built
    return PF_OK;
}

// Spread list: every (key, sample) pair for the tiles a sample's footprint
// touches, stable-sorted by key; then work items with heavy tiles split.
static int build_slist(pf_plan* p, const TapParams& P, int merge, hipStream_t stream, pf_slist& out,
                       const std::vector<uint8_t>& cmask) {
    const int64_t M = p->mtot;
    const int TB = 256;
    const uint64_t nkeys = (uint64_t)(merge ? 1 : p->nframes) * p->ntile_tot;
    const unsigned nbits = bits_for(nkeys + 1);
    DevBuf cnt, off, skeys, svals, skeys2;
    PF_HIP(cnt.alloc(M * 4));
    PF_HIP(off.alloc((M + 1) * 4));
    hipLaunchKernelGGL(k_slist_count, dim3(nblk(M, TB)), dim3(TB), 0, stream, p->taprec, P,
                       cnt.as<uint32_t>());
    PF_HIP(hipGetLastError());
    size_t sbytes = 0;
    DevBuf stmp;
    PF_HIP(rocprim::exclusive_scan(nullptr, sbytes, cnt.as<uint32_t>(), off.as<uint32_t>(), 0u,
                                   (size_t)M, rocprim::plus<uint32_t>(), stream));
    PF_HIP(stmp.alloc(sbytes));
    PF_HIP(rocprim::exclusive_scan(stmp.p, sbytes, cnt.as<uint32_t>(), off.as<uint32_t>(), 0u,
                                   (size_t)M, rocprim::plus<uint32_t>(), stream));
    uint32_t last_off = 0, last_cnt = 0;
    PF_HIP(hipMemcpyAsync(&last_off, off.as<uint32_t>() + (M - 1), 4, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipMemcpyAsync(&last_cnt, cnt.as<uint32_t>() + (M - 1), 4, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));
    const int64_t NE = (int64_t)last_off + last_cnt;
    if (NE >= (1ll << 31)) return pf_fail(PF_ERR_UNSUPPORTED, "spread list exceeds 2^31 entries");
    out.nslist = NE;
    PF_HIP(skeys.alloc(NE * 4));
    PF_HIP(skeys2.alloc(NE * 4));
    PF_HIP(svals.alloc(NE * 4));
    // PF_SLIST_PAD ids past the end: whole 16-entry batches of the last
    // item can be read as one piece
    PF_HIP(hipMalloc(&out.slist, (NE + PF_SLIST_PAD) * 4));
    PF_HIP(hipMemsetAsync(out.slist + NE, 0, PF_SLIST_PAD * 4, stream));
    hipLaunchKernelGGL(k_slist_emit, dim3(nblk(M, TB)), dim3(TB), 0, stream, p->taprec, P, merge,
                       off.as<uint32_t>(), skeys.as<uint32_t>(), svals.as<int32_t>());
    PF_HIP(hipGetLastError());
    size_t rbytes = 0;
    DevBuf rtmp;
    PF_HIP(rocprim::radix_sort_pairs(nullptr, rbytes, skeys.as<uint32_t>(), skeys2.as<uint32_t>(),
                                     svals.as<int32_t>(), out.slist, (size_t)NE, 0, nbits, stream));
    PF_HIP(rtmp.alloc(rbytes));
    PF_HIP(rocprim::radix_sort_pairs(rtmp.p, rbytes, skeys.as<uint32_t>(), skeys2.as<uint32_t>(),
                                     svals.as<int32_t>(), out.slist, (size_t)NE, 0, nbits, stream));
    DevBuf dstart, dend;
    PF_HIP(dstart.alloc(nkeys * 4));
    PF_HIP(dend.alloc(nkeys * 4));
    PF_HIP(hipMemsetAsync(dstart.p, 0, nkeys * 4, stream));
    PF_HIP(hipMemsetAsync(dend.p, 0, nkeys * 4, stream));
    hipLaunchKernelGGL(k_bounds, dim3(nblk(NE, TB)), dim3(TB), 0, stream, skeys2.as<uint32_t>(), NE,
                       dstart.as<int32_t>(), dend.as<int32_t>());
    PF_HIP(hipGetLastError());
    std::vector<int32_t> hs(nkeys), he(nkeys);
    PF_HIP(hipMemcpyAsync(hs.data(), dstart.p, nkeys * 4, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipMemcpyAsync(he.data(), dend.p, nkeys * 4, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));

    // work items: heavy tiles split into chunks that a reduce pass sums in order
    // entries per work item: a chunk is summed serially by one wave, so a
    // tile far heavier than the rest (the k-space centre of radial and
    // spiral trajectories) bounds the kernel unless it is split; splitting
    // tiles of ordinary weight only adds partial-tile traffic.  Split [0]:
    // above four times the mean entries of a non-empty tile, within
    // 128..1024.  Split [1]: also small enough that the list makes at least
    // ~4096 work items (a 2-D grid has few tiles: one wave per tile would
    // leave most of the GPU idle), within 64..1024; the launch picks it for
    // batches whose value chunks do not fill the GPU anyway.  Split [2]: twice
    // split [1]'s item size, for such batches of several value chunks (a
    // 64-value 2-D B0 adjoint measured 0.204 -> 0.191 ms with it, while
    // one-chunk batches measured slower).
    int CH[3] = {p->max_chunk, p->max_chunk, p->max_chunk};
    if (p->max_chunk <= 0) {
        int64_t nonempty = 0;
        for (uint64_t k = 0; k < nkeys; ++k) nonempty += he[k] > hs[k];
        const int64_t heavy = 4 * NE / std::max<int64_t>(1, nonempty);
        const int64_t fill = (NE + 4095) / 4096;
        CH[0] = (int)std::min<int64_t>(1024, std::max<int64_t>(128, heavy));
        CH[1] = (int)std::min<int64_t>(1024, std::max<int64_t>(PF_SPREAD_MINCH, std::min(heavy, fill)));
        CH[2] = std::min(CH[0], 2 * CH[1]);
    }
    // visit tiles in the sample order's Z-order (when it is used), frame by frame
    std::vector<uint64_t> order(nkeys);
    {
        const int msum = P.mbits[0] + P.mbits[1] + P.mbits[2];
        std::vector<std::pair<uint64_t, uint64_t>> tk(p->ntile_tot);
        for (int64_t t = 0; t < p->ntile_tot; ++t) {
            const uint32_t tx = (uint32_t)(t % p->ntiles[2]);
            const uint32_t ty = (uint32_t)((t / p->ntiles[2]) % p->ntiles[1]);
            const uint32_t tz = (uint32_t)(t / ((int64_t)p->ntiles[2] * p->ntiles[1]));
            tk[t] = {msum > 0 ? morton3(tz, ty, tx, P.mbits) : (uint64_t)t, (uint64_t)t};
        }
        std::sort(tk.begin(), tk.end());
        uint64_t q = 0;
        for (uint64_t f = 0; f < nkeys / p->ntile_tot; ++f)
            for (int64_t t = 0; t < p->ntile_tot; ++t) order[q++] = f * p->ntile_tot + tk[t].second;
    }
    for (int m = 0; m < 3; ++m) {
        pf_sitems& S = out.split[m];
        bool shared = false;
        for (int q = 0; q < m && !shared; ++q)
            if (CH[q] == CH[m]) {
                S = out.split[q];  // same arrays
                shared = true;
            }
        if (shared) continue;
        std::vector<int4> items, ritems;
        int slot = 0;
        for (uint64_t k : order) {
            int e0 = hs[k], e1 = he[k];
            int n = e1 - e0;
            if (n <= CH[m]) {
                items.push_back(make_int4((int)k, e0, e1, -1));
            } else {
                int nch = (n + CH[m] - 1) / CH[m];
                ritems.push_back(make_int4((int)k, slot, nch, 0));
                for (int c = 0; c < nch; ++c) {
                    int a = e0 + (int)((int64_t)n * c / nch);
                    int b = e0 + (int)((int64_t)n * (c + 1) / nch);
                    items.push_back(make_int4((int)k, a, b, slot++));
                }
            }
        }
        S.nsitems = (int)items.size();
        S.nritems = (int)ritems.size();
        S.nslots = slot;
        PF_HIP(hipMalloc(&S.sitems, std::max<size_t>(1, items.size()) * sizeof(int4)));
        PF_HIP(hipMemcpy(S.sitems, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
        if (!cmask.empty()) {
            // drop the empty tiles lying wholly in columns no sample covers
            std::vector<int4> live;
            live.reserve(items.size());
            for (const int4& it : items) {
                bool keep = it.z > it.y;
                if (!keep) {
                    const int64_t t = it.x % p->ntile_tot;
                    const int tx = (int)(t % p->ntiles[2]), ty = (int)((t / p->ntiles[2]) % p->ntiles[1]);
                    const int y1 = std::min(p->os[1], (ty + 1) * p->tile[1]);
                    const int x1 = std::min(p->os[2], (tx + 1) * p->tile[2]);
                    for (int y = ty * p->tile[1]; y < y1 && !keep; ++y)
                        for (int x = tx * p->tile[2]; x < x1 && !keep; ++x) keep = cmask[(size_t)y * p->os[2] + x] != 0;
                }
                if (keep) live.push_back(it);
            }
            S.nlsitems = (int)live.size();
            PF_HIP(hipMalloc(&S.lsitems, std::max<size_t>(1, live.size()) * sizeof(int4)));
            PF_HIP(hipMemcpy(S.lsitems, live.data(), live.size() * sizeof(int4), hipMemcpyHostToDevice));
        }
        PF_HIP(hipMalloc(&S.ritems, std::max<size_t>(1, ritems.size()) * sizeof(int4)));
        if (!ritems.empty())
            PF_HIP(hipMemcpy(S.ritems, ritems.data(), ritems.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    return PF_OK;
}

static int build_glist(pf_plan* p, const TapParams& P, hipStream_t stream);

// 3-D column mask: every sample marks the (y, x) grid columns of its
// footprint (plain byte stores of 1, so concurrent marks agree)
__global__ void k_colmask(const float* __restrict__ taprec, int rec, int64_t M, int Gy, int Gx, int Wy, int Wx,
                          uint8_t* __restrict__ mask) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M * Wy) return;
    const int64_t s = j / Wy;
    const int dy = (int)(j % Wy);
    const int* ri = reinterpret_cast<const int*>(taprec + s * rec);
    int y = ri[1] + dy;
    y -= y >= Gy ? Gy : 0;
    for (int dx = 0; dx < Wx; ++dx) {
        int x = ri[2] + dx;
        x -= x >= Gx ? Gx : 0;
        mask[(int64_t)y * Gx + x] = 1;
    }
}

int pf_plan_build(pf_plan* p, const float* coord, hipStream_t stream) {
    TapParams P;
    P.ndim = p->ndim;
    for (int a = 0; a < 3; ++a) {
        P.os[a] = p->os[a];
        P.width[a] = p->width[a];
        P.beta[a] = p->beta[a];
        float G = (float)p->os[a], N = (float)p->shape[a];
        P.sc[a] = G / N;
        P.sh[a] = (float)(p->os[a] / 2);
        P.tile[a] = p->tile[a];
        P.ntiles[a] = p->ntiles[a];
        P.woff[a] = p->woff[a];
    }
    P.npts = p->npts;
    P.mtot = p->mtot;
    P.ntile_tot = p->ntile_tot;
    P.rec = p->rec;
    P.foff = p->foff;
    P.nframes = p->nframes;
    int msum = 0, isum = 0;
    for (int a = 0; a < 3; ++a) {
        P.mbits[a] = 0;
        while ((1 << P.mbits[a]) < p->ntiles[a]) ++P.mbits[a];
        msum += P.mbits[a];
        P.ibits[a] = 0;
        while ((1 << P.ibits[a]) < p->tile[a]) ++P.ibits[a];
        isum += P.ibits[a];
    }
    // Z-order keys need frame << (msum + isum) to fit 32 bits: drop the
    // in-tile order first, then fall back to row-major tile order
    if ((uint64_t)p->nframes << (msum + isum) >= (1ull << 32)) P.ibits[0] = P.ibits[1] = P.ibits[2] = 0, isum = 0;
    if (msum == 0 || (uint64_t)p->nframes << msum >= (1ull << 32))
        P.mbits[0] = P.mbits[1] = P.mbits[2] = 0, msum = 0, P.ibits[0] = P.ibits[1] = P.ibits[2] = 0, isum = 0;
    const int64_t M = p->mtot;
    const int TB = 256;

    DevBuf keys, keys2, ids, rec0, cos0, tmp;
    PF_HIP(keys.alloc(M * 4));
    PF_HIP(keys2.alloc(M * 4));
    PF_HIP(ids.alloc(M * 4));
    PF_HIP(rec0.alloc(M * p->rec * 4));
    PF_HIP(cos0.alloc(M * 3 * 4));
    PF_HIP(hipMalloc(&p->perm, M * 4));
    PF_HIP(hipMalloc(&p->iperm, M * 4));
    PF_HIP(hipMalloc(&p->taprec, M * p->rec * 4));
    PF_HIP(hipMalloc(&p->cos_sorted, M * 3 * 4));
    PF_HIP(hipMemsetAsync(rec0.p, 0, M * p->rec * 4, stream));

    hipLaunchKernelGGL(k_taps, dim3(nblk(M, TB)), dim3(TB), 0, stream, coord, P, keys.as<uint32_t>(),
                       ids.as<int32_t>(), rec0.as<float>(), cos0.as<float>());
    PF_HIP(hipGetLastError());

    const uint64_t nkeys = msum > 0 ? (uint64_t)p->nframes << (msum + isum) : (uint64_t)p->nframes * p->ntile_tot;
    unsigned nbits = bits_for(nkeys + 1);
    size_t tbytes = 0;
    PF_HIP(rocprim::radix_sort_pairs(nullptr, tbytes, keys.as<uint32_t>(), keys2.as<uint32_t>(),
                                     ids.as<int32_t>(), p->perm, (size_t)M, 0, nbits, stream));
    PF_HIP(tmp.alloc(tbytes));
    PF_HIP(rocprim::radix_sort_pairs(tmp.p, tbytes, keys.as<uint32_t>(), keys2.as<uint32_t>(),
                                     ids.as<int32_t>(), p->perm, (size_t)M, 0, nbits, stream));
    hipLaunchKernelGGL(k_permute, dim3(nblk(M, TB)), dim3(TB), 0, stream, p->perm, rec0.as<float>(),
                       p->taprec, cos0.as<float>(), p->cos_sorted, p->iperm, M, p->rec);
    PF_HIP(hipGetLastError());

    std::vector<uint8_t> cmask;
    if (p->ndim == 3 && PF_COLMASK) {
        const int64_t ncol = (int64_t)p->os[1] * p->os[2];
        PF_HIP(hipMalloc(&p->colmask, ncol));
        PF_HIP(hipMemsetAsync(p->colmask, 0, ncol, stream));
        hipLaunchKernelGGL(k_colmask, dim3(nblk(M * p->width[1], TB)), dim3(TB), 0, stream, p->taprec, p->rec, M,
                           p->os[1], p->os[2], p->width[1], p->width[2], p->colmask);
        PF_HIP(hipGetLastError());
        cmask.resize(ncol);
        PF_HIP(hipMemcpyAsync(cmask.data(), p->colmask, ncol, hipMemcpyDeviceToHost, stream));
    }
    PF_HIP(hipStreamSynchronize(stream));
    int rc = build_slist(p, P, 0, stream, p->sl[0], cmask);
    if (rc != PF_OK) return rc;
    if ((rc = build_glist(p, P, stream)) != PF_OK) return rc;
    if (p->nframes > 1) return build_slist(p, P, 1, stream, p->sl[1], cmask);
    p->sl[1] = p->sl[0];
    return PF_OK;
}

// Lane order of one window's samples.  Every tap adds the same offset to
// all lanes' first-tap LDS slots (16-byte points), so the bank pattern of a
// whole tap loop is that of the first taps.  Measured on gfx950 (64 lanes of
// ds_read_b128): a 16-lane group is free of bank conflicts when lane j of
// the group reads a slot congruent to j mod 16, or when lanes share one
// address (a broadcast); any other mix of the 16 residues costs about 1.5x.
// So a window's samples fill groups of 16 by residue: lane j takes the next
// sample of residue class j, as long as every class has samples, which
// keeps groups conflict-free.  The slots a class cannot fill are then given
// to the leftover samples of the fuller classes, first to a group that
// already holds the same first-tap position (a broadcast), else to the
// first free slot; every lane slot but the end of the window's last block
// is used.
static void balance_lanes(const int32_t* ids, const int32_t* base, int n, std::vector<int32_t>& lanes) {
    std::vector<std::pair<int32_t, int32_t>> pb(n);  // (first-tap position, sample), by position
    for (int i = 0; i < n; ++i) pb[i] = {base[i], ids[i]};
    std::stable_sort(pb.begin(), pb.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int> cls[16];  // indices into pb, by residue, positions in order
    for (int i = 0; i < n; ++i) cls[pb[i].first & 15].push_back(i);
    const int ng = (n + 15) / 16;
    std::vector<int> slot((size_t)ng * 16, -1);  // index into pb, or -1
    size_t head[16] = {0};
    for (int g = 0; g < ng; ++g)
        for (int r = 0; r < 16; ++r)
            if (head[r] < cls[r].size()) slot[(size_t)g * 16 + r] = cls[r][head[r]++];
    // leftovers: a class's samples are in position order, so the last group
    // that took a sample of the same position is the previous sample's
    size_t free_scan = 0;
    for (int r = 0; r < 16; ++r) {
        int prev_group = head[r] > 0 ? ng - 1 : -1;  // group of cls[r][head[r] - 1]
        for (; head[r] < cls[r].size(); ++head[r]) {
            const int i = cls[r][head[r]];
            int at = -1;
            if (prev_group >= 0 && pb[cls[r][head[r] - 1]].first == pb[i].first)
                for (int q = 0; q < 16 && at < 0; ++q)
                    if (slot[(size_t)prev_group * 16 + q] < 0) at = prev_group * 16 + q;
            if (at < 0) {
                while (slot[free_scan] >= 0) ++free_scan;
                at = (int)free_scan;
            }
            slot[at] = i;
            prev_group = at / 16;
        }
    }
    for (int v : slot) lanes.push_back(v >= 0 ? pb[v].second : -1);
}

// Gather lists of the LDS-staged gather (see pf_glist); left empty (the
// direct-read gather runs) for plans it does not cover.
static int build_glist(pf_plan* p, const TapParams& P, hipStream_t stream) {
    const int W = p->width[0];
    if (p->ndim != 3 || W < 3 || p->width[1] != W || p->width[2] != W) return PF_OK;
    const int dz = pf_gather_dz(W);
    pf_glist& L = p->gl;
    L.ncol[0] = (p->os[1] + PF_GCOL - 1) / PF_GCOL;
    L.ncol[1] = (p->os[2] + PF_GCOL - 1) / PF_GCOL;
    L.nzw = (p->os[0] + dz - 1) / dz;
    const uint64_t nkeys = (uint64_t)p->nframes * L.ncol[0] * L.ncol[1] * L.nzw;
    if (nkeys >= (1ull << 31)) return PF_OK;
    const int64_t M = p->mtot;
    const int TB = 256;
    const unsigned nbits = bits_for(nkeys + 1);
    DevBuf keys, keys2, vals, vals2, bvals, bvals2, tmp;
    PF_HIP(keys.alloc(M * 4));
    PF_HIP(keys2.alloc(M * 4));
    PF_HIP(vals.alloc(M * 4));
    PF_HIP(vals2.alloc(M * 4));
    PF_HIP(bvals.alloc(M * 4));
    PF_HIP(bvals2.alloc(M * 4));
    hipLaunchKernelGGL(k_gather_keys, dim3(nblk(M, TB)), dim3(TB), 0, stream, p->taprec, P, L.ncol[0], L.ncol[1], L.nzw,
                       dz, keys.as<uint32_t>(), vals.as<int32_t>(), bvals.as<int32_t>());
    PF_HIP(hipGetLastError());
    // sort (key, sample) pairs, then carry the first-tap positions along
    // through the sorted sample ids (vals2[i] is the sorted index t)
    size_t rbytes = 0;
    PF_HIP(rocprim::radix_sort_pairs(nullptr, rbytes, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<int32_t>(),
                                     vals2.as<int32_t>(), (size_t)M, 0, nbits, stream));
    PF_HIP(tmp.alloc(rbytes));
    PF_HIP(rocprim::radix_sort_pairs(tmp.p, rbytes, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<int32_t>(),
                                     vals2.as<int32_t>(), (size_t)M, 0, nbits, stream));
    std::vector<uint32_t> hk(M);
    std::vector<int32_t> hid(M), hb(M);
    PF_HIP(hipMemcpyAsync(hk.data(), keys2.p, M * 4, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipMemcpyAsync(hid.data(), vals2.p, M * 4, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipMemcpyAsync(hb.data(), bvals.p, M * 4, hipMemcpyDeviceToHost, stream));  // indexed by t
    PF_HIP(hipStreamSynchronize(stream));
    std::vector<int32_t> lanes, hbs;
    std::vector<int4> items;
    lanes.reserve((size_t)M + M / 4);
    for (int64_t e0 = 0; e0 < M;) {
        int64_t e1 = e0;
        while (e1 < M && hk[e1] == hk[e0]) ++e1;
        const int n = (int)(e1 - e0);
        hbs.resize(n);
        for (int i = 0; i < n; ++i) hbs[i] = hb[hid[e0 + i]];
        const size_t l0 = lanes.size();
        balance_lanes(hid.data() + e0, hbs.data(), n, lanes);
        // items of PF_GCOL_MAXS lanes (whole groups); the last one padded
        const size_t nl = lanes.size() - l0;
        const size_t nit = (nl + PF_GCOL_MAXS - 1) / PF_GCOL_MAXS;
        lanes.resize(l0 + nit * PF_GCOL_MAXS, -1);
        for (size_t k = 0; k < nit; ++k)
            items.push_back(make_int4((int)hk[e0], (int)(l0 + k * PF_GCOL_MAXS), 0, 0));
        e0 = e1;
    }
    if ((int64_t)lanes.size() >= (1ll << 31)) return PF_OK;
    int64_t placed = 0;
    for (int32_t v : lanes) placed += v >= 0;
    if (placed != M) return pf_fail(PF_ERR_ARG, "gather lane assignment lost samples (internal error)");
    L.nitems = (int)items.size();
    PF_HIP(hipMalloc(&L.ids, std::max<size_t>(1, lanes.size()) * sizeof(int32_t)));
    PF_HIP(hipMemcpy(L.ids, lanes.data(), lanes.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    PF_HIP(hipMalloc(&L.items, std::max<size_t>(1, items.size()) * sizeof(int4)));
    if (!items.empty())
        PF_HIP(hipMemcpy(L.items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
    return PF_OK;
}

void pf_plan_free(pf_plan* p) {
    if (!p) return;
    void* ptrs[] = {p->perm, p->iperm, p->taprec, p->cos_sorted, p->apod, p->foff, p->gl.ids, p->gl.items, p->colmask};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    const bool alias = p->sl[1].slist == p->sl[0].slist;
    for (int m = 0; m < (alias ? 1 : 2); ++m) {
        const pf_slist& L = p->sl[m];
        if (L.slist) (void)hipFree(L.slist);
        for (int k = 0; k < 3; ++k) {
            bool dup = false;  // arrays shared with an earlier split
            for (int q = 0; q < k; ++q) dup = dup || L.split[q].sitems == L.split[k].sitems;
            if (dup) continue;
            if (L.split[k].sitems) (void)hipFree(L.split[k].sitems);
            if (L.split[k].ritems) (void)hipFree(L.split[k].ritems);
            if (L.split[k].lsitems) (void)hipFree(L.split[k].lsitems);
        }
    }
    if (p->fft) pf_fft_destroy(p->fft);
    if (p->fftk) pf_fftk_destroy(p->fftk);
    delete p;
}
