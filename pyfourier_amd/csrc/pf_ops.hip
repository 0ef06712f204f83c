// Operator kernels of the NUFFT engine (gfx950, wave64).
//
//   deapod+pad   image (C, nimg, N)  -> grid (nfg, Cb, G, R, Cv)
//                _subroutines/_apod.py:16-46 then _resize.py:16-42 and the
//                ifftshift of _fft.py:72-79, folded into one index map
//   gather       grid -> samples (sorted, coil innermost)
//                _cart2noncart/_3D.py:13-51, subspace twin _3D_subspace.py
//   spread       samples -> grid, deterministic tile accumulation in LDS
//                (no global float atomics), _noncart2cart/_3D.py:13-51
//   crop+deapod  grid -> image, _nufft/_nufft.py:271-282
//   sort/unsort  user sample order (C, F, npts) <-> tile-sorted (Mtot, C),
//                with the density weight and output scaling fused
#include <algorithm>

#include "pf_gcol.h"
#include "pf_internal.h"

namespace {

constexpr int TB = 256;
#ifndef PF_GATHER_TB
#define PF_GATHER_TB 256
#endif
constexpr int GTB = PF_GATHER_TB;  // gather block: samples of a block share the CU's L1
#ifndef PF_GATHER_COL
#define PF_GATHER_COL 1  // LDS-staged gather where it applies
#endif

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

struct GeoParams {
    int os[3];
    int shape[3];
    int width[3];
    int tile[3];
    int ntiles[3];
    int64_t ntile_tot;
    int64_t gtot;
    int64_t ntot;
    int64_t npts;
    int rec;
    int woff[3];
};

GeoParams geo(const pf_plan* p) {
    GeoParams g;
    for (int a = 0; a < 3; ++a) {
        g.os[a] = p->os[a];
        g.shape[a] = p->shape[a];
        g.width[a] = p->width[a];
        g.tile[a] = p->tile[a];
        g.ntiles[a] = p->ntiles[a];
        g.woff[a] = p->woff[a];
    }
    g.ntile_tot = p->ntile_tot;
    g.gtot = pf_gtot(p);
    g.ntot = pf_ntot(p);
    g.npts = p->npts;
    g.rec = p->rec;
    return g;
}

struct BatchParams {
    int C, Cv, Cb, R, L, K, nfg, nimg, nt;
    const float2* basis;
    const float2* b0t;
    const float2* b0s;
    int dma;  // spread values copied global -> LDS (set by the spread launcher)
};

BatchParams bparams(const pf_batch& b) {
    BatchParams q{b.C, b.Cv, b.Cb, b.R, b.L, b.K, b.nfg, b.nimg, b.nt, b.basis, b.b0t, b.b0s, 0};
    return q;
}

// Coefficient of reduction entry r = (l, k) for a sample of frame f and
// readout index t: Bt[l, t] * basis[f, k] (either factor 1 when absent).
__device__ __forceinline__ float2 sample_coef(const BatchParams& B, int r, int f, int64_t t) {
    float2 c = make_float2(1.0f, 0.0f);
    int l = r / B.K, k = r % B.K;
    if (B.b0t) c = B.b0t[(int64_t)l * B.nt + t];
    if (B.basis) c = cmul(c, B.basis[(int64_t)f * B.K + k]);
    return c;
}

// -------------------------------------------------------------------------
// deapod + zero-pad + ifftshift: one thread per grid value
// -------------------------------------------------------------------------
__global__ void k_deapod_pad(GeoParams g, BatchParams B, const float* __restrict__ apod,
                             const float2* __restrict__ img, float2* __restrict__ grid) {
    const int RC = B.R * B.Cv;
    const int64_t total = (int64_t)B.nfg * B.Cb * g.gtot * RC;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int cv = (int)(e % B.Cv);
        int r = (int)((e / B.Cv) % B.R);
        int64_t pt = (e / RC) % g.gtot;
        int64_t gc = e / (RC * g.gtot);  // g*Cb + cb
        int cb = (int)(gc % B.Cb);
        int fg = (int)(gc / B.Cb);
        int nx = (int)(pt % g.os[2]);
        int ny = (int)((pt / g.os[2]) % g.os[1]);
        int nz = (int)(pt / ((int64_t)g.os[2] * g.os[1]));
        int qx = (nx + g.shape[2] / 2) % g.os[2];
        int qy = (ny + g.shape[1] / 2) % g.os[1];
        int qz = (nz + g.shape[0] / 2) % g.os[0];
        float2 v = make_float2(0.f, 0.f);
        if (qx < g.shape[2] && qy < g.shape[1] && qz < g.shape[0]) {
            int64_t p = ((int64_t)qz * g.shape[1] + qy) * g.shape[2] + qx;
            int l = r / B.K, k = r % B.K;
            int ii = B.nfg > 1 ? fg : (B.nimg > 1 ? k : 0);
            int c = cb * B.Cv + cv;
            v = img[((int64_t)c * B.nimg + ii) * g.ntot + p];
            float a = apod[qz] * apod[g.shape[0] + qy] * apod[g.shape[0] + g.shape[1] + qx];
            v.x *= a;
            v.y *= a;
            if (B.b0s) v = cmul(v, B.b0s[(int64_t)l * g.ntot + p]);
        }
        grid[e] = v;
    }
}

// -------------------------------------------------------------------------
// crop + fftshift + deapod (+ conj B0 maps, summed over segments)
// -------------------------------------------------------------------------
__global__ void k_crop_deapod(GeoParams g, BatchParams B, const float* __restrict__ apod,
                              const float2* __restrict__ grid, float2* __restrict__ img, float factor) {
    const int RC = B.R * B.Cv;
    const int64_t total = (int64_t)B.C * B.nimg * g.ntot;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = e % g.ntot;
        int ii = (int)((e / g.ntot) % B.nimg);
        int c = (int)(e / (g.ntot * B.nimg));
        int qx = (int)(p % g.shape[2]);
        int qy = (int)((p / g.shape[2]) % g.shape[1]);
        int qz = (int)(p / ((int64_t)g.shape[2] * g.shape[1]));
        int nx = (qx - g.shape[2] / 2 + g.os[2]) % g.os[2];
        int ny = (qy - g.shape[1] / 2 + g.os[1]) % g.os[1];
        int nz = (qz - g.shape[0] / 2 + g.os[0]) % g.os[0];
        int64_t pt = ((int64_t)nz * g.os[1] + ny) * g.os[2] + nx;
        int fg = B.nfg > 1 ? ii : 0;
        int k = B.nfg > 1 ? 0 : ii;
        int cb = c / B.Cv, cv = c % B.Cv;
        const float2* base = grid + (((int64_t)fg * B.Cb + cb) * g.gtot + pt) * RC + cv;
        float2 acc = make_float2(0.f, 0.f);
        for (int l = 0; l < B.L; ++l) {
            float2 v = base[(l * B.K + k) * B.Cv];
            if (B.b0s) v = cmulc(v, B.b0s[(int64_t)l * g.ntot + p]);
            acc.x += v.x;
            acc.y += v.y;
        }
        float a = factor * apod[qz] * apod[g.shape[0] + qy] * apod[g.shape[0] + g.shape[1] + qx];
        img[e] = make_float2(acc.x * a, acc.y * a);
    }
}

// -------------------------------------------------------------------------
// gather: one lane per (sorted sample, CPL coils)
// -------------------------------------------------------------------------
template <int CPL>
__global__ void __launch_bounds__(TB) k_gather(GeoParams g, BatchParams B, int64_t mtot,
                                               const float* __restrict__ taprec,
                                               const int32_t* __restrict__ perm,
                                               const float2* __restrict__ grid,
                                               float2* __restrict__ ys) {
    const int LS = B.C / CPL;  // lanes per sample
    unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    int64_t t = (int64_t)bid * blockDim.x + threadIdx.x;
    int64_t s = t / LS;
    if (s >= mtot) return;
    int lane = (int)(t % LS);
    int c0 = lane * CPL;
    int cb = c0 / B.Cv, cv = c0 % B.Cv;
    const float* rec = taprec + s * g.rec;
    const int* ri = reinterpret_cast<const int*>(rec);
    int f = ri[3];
    int fg = B.nfg > 1 ? f : 0;
    int64_t tread = 0;
    if (B.b0t) tread = ((int64_t)perm[s] % g.npts) % B.nt;
    int iz[PF_WMAX], iy[PF_WMAX], ix[PF_WMAX];
    float wz[PF_WMAX], wy[PF_WMAX], wx[PF_WMAX];
    const int Wz = g.width[0], Wy = g.width[1], Wx = g.width[2];
    for (int k = 0; k < Wz; ++k) {
        int v = ri[0] + k;
        iz[k] = v >= g.os[0] ? v - g.os[0] : v;
        wz[k] = rec[g.woff[0] + k];
    }
    for (int k = 0; k < Wy; ++k) {
        int v = ri[1] + k;
        iy[k] = v >= g.os[1] ? v - g.os[1] : v;
        wy[k] = rec[g.woff[1] + k];
    }
    for (int k = 0; k < Wx; ++k) {
        int v = ri[2] + k;
        ix[k] = v >= g.os[2] ? v - g.os[2] : v;
        wx[k] = rec[g.woff[2] + k];
    }
    const int RC = B.R * B.Cv;
    const float2* slab = grid + ((int64_t)fg * B.Cb + cb) * g.gtot * RC + cv;
    float2 out[CPL];
    for (int q = 0; q < CPL; ++q) out[q] = make_float2(0.f, 0.f);
    for (int r = 0; r < B.R; ++r) {
        float2 acc[CPL];
        for (int q = 0; q < CPL; ++q) acc[q] = make_float2(0.f, 0.f);
        const float2* sr = slab + r * B.Cv;
        for (int a = 0; a < Wz; ++a) {
            for (int b = 0; b < Wy; ++b) {
                float wzy = wz[a] * wy[b];
                int64_t row = ((int64_t)iz[a] * g.os[1] + iy[b]) * g.os[2];
                for (int c = 0; c < Wx; ++c) {
                    float w = wzy * wx[c];
                    const float2* src = sr + (row + ix[c]) * RC;
                    if constexpr (CPL == 2) {
                        float4 v = *reinterpret_cast<const float4*>(src);
                        acc[0].x += w * v.x;
                        acc[0].y += w * v.y;
                        acc[CPL - 1].x += w * v.z;
                        acc[CPL - 1].y += w * v.w;
                    } else {
                        float2 v = *src;
                        acc[0].x += w * v.x;
                        acc[0].y += w * v.y;
                    }
                }
            }
        }
        if (B.R == 1 && !B.basis && !B.b0t) {
            for (int q = 0; q < CPL; ++q) out[q] = acc[q];
        } else {
            float2 cf = sample_coef(B, r, f, tread);
            for (int q = 0; q < CPL; ++q) {
                float2 m = cmul(acc[q], cf);
                out[q].x += m.x;
                out[q].y += m.y;
            }
        }
    }
    float2* dst = ys + s * B.C + c0;
    if constexpr (CPL == 2)
        *reinterpret_cast<float4*>(dst) = make_float4(out[0].x, out[0].y, out[CPL - 1].x, out[CPL - 1].y);
    else
        dst[0] = out[0];
}

// Gather with compile-time y and x widths: the WY x WX taps of a z row are
// fully unrolled (all their loads in flight together, tap indices and weights
// in registers by constant index), z stays a loop whose weight is read from
// the record.  In-slab offsets are 32-bit (the host checks G * R * Cv < 2^31).
// DEEP (3-D plans, several z rows per sample): the compiler may plan for as
// few as 2 waves per SIMD and keep more of a sample's tap loads in flight
// per wave (measured 13 % faster on 128^3 x 32 coils at W=6; 2-D plans, one z
// row per sample, are faster at the default budget)
template <int CPL, int WY, int WX, bool DEEP = false>
__global__ void __launch_bounds__(GTB) __attribute__((amdgpu_waves_per_eu(1, DEEP ? 2 : 8))) k_gather_w(GeoParams g, BatchParams B, int64_t mtot,
                                                 const float* __restrict__ taprec,
                                                 const int32_t* __restrict__ perm,
                                                 const float2* __restrict__ grid, float2* __restrict__ ys) {
    const int LS = B.C / CPL;  // lanes per sample
    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t t = (int64_t)bid * blockDim.x + threadIdx.x;
    const int64_t s = t / LS;
    if (s >= mtot) return;
    const int lane = (int)(t % LS);
    const int c0 = lane * CPL;
    const int cb = c0 / B.Cv, cv = c0 % B.Cv;
    const float* rec = taprec + s * g.rec;
    const int* ri = reinterpret_cast<const int*>(rec);
    const int f = ri[3];
    const int fg = B.nfg > 1 ? f : 0;
    const int64_t tread = B.b0t ? ((int64_t)perm[s] % g.npts) % B.nt : 0;
    const int RC = B.R * B.Cv;
    int yo[WY], xo[WX];
    float wy[WY], wx[WX];
#pragma unroll
    for (int k = 0; k < WY; ++k) {
        int v = ri[1] + k;
        v -= v >= g.os[1] ? g.os[1] : 0;
        yo[k] = v * g.os[2] * RC;
        wy[k] = rec[g.woff[1] + k];
    }
#pragma unroll
    for (int k = 0; k < WX; ++k) {
        int v = ri[2] + k;
        v -= v >= g.os[2] ? g.os[2] : 0;
        xo[k] = v * RC;
        wx[k] = rec[g.woff[2] + k];
    }
    const int zstride = g.os[1] * g.os[2] * RC;
    const float2* slab = grid + ((int64_t)fg * B.Cb + cb) * g.gtot * RC + cv;
    float2 out[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) out[q] = make_float2(0.f, 0.f);
    for (int r = 0; r < B.R; ++r) {
        float2 acc[CPL];
#pragma unroll
        for (int q = 0; q < CPL; ++q) acc[q] = make_float2(0.f, 0.f);
        const float2* sr = slab + r * B.Cv;
        for (int a = 0; a < g.width[0]; ++a) {
            int z = ri[0] + a;
            z -= z >= g.os[0] ? g.os[0] : 0;
            const float wz = rec[g.woff[0] + a];
            const float2* pz = sr + z * zstride;
#pragma unroll
            for (int b = 0; b < WY; ++b) {
                const float wzy = wz * wy[b];
                const float2* py = pz + yo[b];
#pragma unroll
                for (int c = 0; c < WX; ++c) {
                    const float w = wzy * wx[c];
                    if constexpr (CPL == 2) {
                        const float4 v = *reinterpret_cast<const float4*>(py + xo[c]);
                        acc[0].x += w * v.x;
                        acc[0].y += w * v.y;
                        acc[CPL - 1].x += w * v.z;
                        acc[CPL - 1].y += w * v.w;
                    } else {
                        const float2 v = py[xo[c]];
                        acc[0].x += w * v.x;
                        acc[0].y += w * v.y;
                    }
                }
            }
        }
        if (B.R == 1 && !B.basis && !B.b0t) {
#pragma unroll
            for (int q = 0; q < CPL; ++q) out[q] = acc[q];
        } else {
            const float2 cf = sample_coef(B, r, f, tread);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const float2 m = cmul(acc[q], cf);
                out[q].x += m.x;
                out[q].y += m.y;
            }
        }
    }
    float2* dst = ys + s * B.C + c0;
    if constexpr (CPL == 2)
        *reinterpret_cast<float4*>(dst) = make_float4(out[0].x, out[0].y, out[CPL - 1].x, out[CPL - 1].y);
    else
        dst[0] = out[0];
}

// -------------------------------------------------------------------------
// LDS-staged gather (3-D plans, one kernel width W on all axes, plain values:
// no basis, no field map; an even batch C).  The forward grid of such a
// batch is laid out coil-pair major, (nfg, C/2, Gz, Gy, Gx, 2), so one coil
// pair of a window row is one contiguous run.
// One block per work item of the plan's gather lists: 256 lanes holding
// samples whose first taps all lie in one window, a PF_GCOL x PF_GCOL column
// in (y, x) times DZ consecutive z planes (lane order chosen at plan time so
// that the 16 lanes of an LDS pass mostly read distinct banks).  For each
// coil pair the block copies the window plus the kernel halo, (DZ + W - 1)
// planes of (PF_GCOL + W - 1)^2 grid points of 16 bytes, from that pair's
// grid into LDS with global -> LDS copies, and every lane reads its sample's
// W^3 taps from LDS.  Two blocks share a CU, so one block's copies overlap
// the other's taps.  The direct gather is bound by the L1 request path (64
// B/clk/CU); conflict-free 16-byte LDS reads run ~2.9x faster, and a staged
// point serves on average ~15 samples on the 3-D radial headline, so the
// copies cost far less than the taps they replace.  A sample's tap record is
// read once; its weights stay in registers across the coil pairs.
// -------------------------------------------------------------------------
// one sample's two-coil value from the staged window: per x tap c the sum
// over (z, y) rows of wz*wy*v, then the x weights once at the end
template <int W>
__device__ __forceinline__ float4 gcol_taps(const float4* win, int base, const float (&wz)[W], const float (&wy)[W],
                                            const float (&wx)[W]) {
    using S = GColShape<W>;
    float4 ax[W];
#pragma unroll
    for (int c = 0; c < W; ++c) ax[c] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const float4* row = win + base + a * S::PP + b * S::PX;
            const float w = wz[a] * wy[b];
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const float4 v = row[c];
                ax[c].x += w * v.x;
                ax[c].y += w * v.y;
                ax[c].z += w * v.z;
                ax[c].w += w * v.w;
            }
        }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int c = 0; c < W; ++c) {
        acc.x += wx[c] * ax[c].x;
        acc.y += wx[c] * ax[c].y;
        acc.z += wx[c] * ax[c].z;
        acc.w += wx[c] * ax[c].w;
    }
    return acc;
}

template <int W>
__global__ void __launch_bounds__(GCOL_TB, 2)
    k_gather_col(GeoParams g, BatchParams B, const int4* __restrict__ items, const int32_t* __restrict__ gids,
                 const float* __restrict__ taprec, const float2* __restrict__ grid, float2* __restrict__ ys, int ncy,
                 int ncx, int nzw) {
    using S = GColShape<W>;
    __shared__ float4 win[S::NPTS];  // staged window, one coil pair
    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int4 it = items[bid];
    const int zw = it.x % nzw;
    const int cx = (it.x / nzw) % ncx;
    const int cy = (it.x / (nzw * ncx)) % ncy;
    const int f = it.x / (nzw * ncx * ncy);
    const int z0 = zw * S::DZ, y0 = cy * PF_GCOL, x0 = cx * PF_GCOL;
    const int ncp = B.C / 2;  // Cb = C / 2 grids of coil pairs (Cv = 2, R = 1)
    const float4* pairs = reinterpret_cast<const float4*>(grid) + (int64_t)(B.nfg > 1 ? f : 0) * B.Cb * g.gtot;
    const int tid = threadIdx.x, wave = tid >> 6;

    // this thread's staged points p = tid, tid + 256, ...: their grid point
    // index (periodic wrap), -1 past the window
    int goff[S::NK];
#pragma unroll
    for (int k = 0; k < S::NK; ++k) {
        const int p = k * GCOL_TB + tid;
        const int pl = p / S::PP, rem = p - pl * S::PP;
        const int py = rem / S::PX, px = rem - py * S::PX;
        const int z = (z0 + pl) % g.os[0], y = (y0 + py) % g.os[1], x = (x0 + px) % g.os[2];
        goff[k] = p < S::NPTS ? (z * g.os[1] + y) * g.os[2] + x : -1;
    }

    // this lane's sample (-1: an unused lane at the end of the window's
    // last block): its first tap inside the window and its weights
    const int t0 = gids[it.y + tid];
    const bool act = t0 >= 0;
    const int t = act ? t0 : 0;
    int base = 0;
    float wz[W], wy[W], wx[W];
#pragma unroll
    for (int k = 0; k < W; ++k) wz[k] = wy[k] = wx[k] = 0.f;
    if (act) {
        const float* rec = taprec + (int64_t)t * g.rec;
        const int* ri = reinterpret_cast<const int*>(rec);
        base = ((ri[0] - z0) * S::PX + (ri[1] - y0)) * S::PX + (ri[2] - x0);
#pragma unroll
        for (int k = 0; k < W; ++k) {
            wz[k] = rec[g.woff[0] + k];
            wy[k] = rec[g.woff[1] + k];
            wx[k] = rec[g.woff[2] + k];
        }
    }
    float4* out = reinterpret_cast<float4*>(ys + (int64_t)t * B.C);

    // the values of GCOL_OUT consecutive pairs are kept and stored together:
    // one 64-byte piece of the sample's row instead of four 16-byte pieces
    // written far apart in time (each a partial cache-line write to HBM)
    float4 o[GCOL_OUT];
    for (int cp = 0; cp < ncp; ++cp) {
        // window -> LDS: lane i of wave w, copy k lands at point k*256 + w*64 + i
        const float4* pair = pairs + (int64_t)cp * g.gtot;
#pragma unroll
        for (int k = 0; k < S::NK; ++k)
            if (goff[k] >= 0) gcol_copy16(pair + goff[k], win + k * GCOL_TB + wave * 64);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (act) {
            const float4 r = gcol_taps<W>(win, base, wz, wy, wx);
            const int q = cp % GCOL_OUT;
#pragma unroll
            for (int i = 0; i < GCOL_OUT; ++i)
                if (i == q) o[i] = r;
            if (q == GCOL_OUT - 1 || cp == ncp - 1) {
#pragma unroll
                for (int i = 0; i < GCOL_OUT; ++i)
                    if (i <= q) out[cp - q + i] = o[i];
            }
        }
        __syncthreads();  // the next pair's copies overwrite the window
    }
}

// -------------------------------------------------------------------------
// spread: one block per (tile work item, value chunk).  A value chunk is a
// range of VC entries of the flat per-point vector (Cb, R, Cv); the waves of
// the block own disjoint parts of it.  Samples are staged into LDS in batches
// of SB (tile-clipped tap box, in-tile weights, values times conj of the
// subspace/B0 coefficient); every wave then walks the batch in order, its
// lanes taking distinct (in-tile tap, value) pairs, so no two lanes add to
// the same LDS word in one instruction and per-wave program order fixes the
// summation order: run-to-run bit-identical with no global atomics.
// The accumulator rows are padded to an odd number of words (no LDS bank
// conflicts between neighbouring taps).
// -------------------------------------------------------------------------
constexpr int SPREAD_ACC_BYTES = 64 * 1024;
constexpr int SB = 32;       // samples staged per batch
constexpr int NWV = TB / 64; // waves per block

// floor(t / c) for 0 <= t < 4096, 1 <= c <= 64, with m = ceil(65536 / c)
__device__ __forceinline__ int fdiv(int t, int m) { return (t * m) >> 16; }

struct SpreadLds {
    int VC, PS, tpts;
    __device__ float* acc(float* base) const { return base; }
};

__global__ void __launch_bounds__(TB) k_spread(GeoParams g, BatchParams B, int VC, const int4* __restrict__ items,
                                               const int32_t* __restrict__ slist,
                                               const float* __restrict__ taprec,
                                               const int32_t* __restrict__ perm,
                                               const float2* __restrict__ ys, float2* __restrict__ grid,
                                               float2* __restrict__ partial) {
    extern __shared__ float lds[];
    const int RC = B.R * B.Cv;
    const int V = B.Cb * RC;
    const int nvch = (V + VC - 1) / VC;
    const int PS = 2 * VC + 1;  // padded words per accumulator point
    const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
    const int tpts = Tx * Ty * Tz;
    float* acc = lds;                                        // tpts * PS
    int* sbox = reinterpret_cast<int*>(acc + tpts * PS);     // SB * 8
    float* sw = reinterpret_cast<float*>(sbox + SB * 8);     // SB * 3 * PF_WMAX
    float2* sy = reinterpret_cast<float2*>(sw + SB * 3 * PF_WMAX + (SB * 3 * PF_WMAX) % 2);  // SB * VC

    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int4 it = items[bid / nvch];
    const int vch = bid % nvch;
    const int vbase = vch * VC;
    const int vcnt = min(VC, V - vbase);
    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tx = (int)(tile % g.ntiles[2]);
    const int ty = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int o[3] = {tz * Tz, ty * Ty, tx * Tx};
    const int e[3] = {min(Tz, g.os[0] - o[0]), min(Ty, g.os[1] - o[1]), min(Tx, g.os[2] - o[2])};

    for (int i = threadIdx.x; i < tpts * PS; i += TB) acc[i] = 0.f;

    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int VW = (vcnt + NWV - 1) / NWV;
    const int vw0 = min(vcnt, wave * VW);
    const int vwn = min(vcnt, vw0 + VW) - vw0;  // values of this wave
    const int mvw = vwn > 0 ? (65536 + vwn - 1) / vwn : 1;

    for (int eb = it.y; eb < it.z; eb += SB) {
        const int nb = min(SB, it.z - eb);
        __syncthreads();  // previous batch fully consumed (and accumulator zeroed)
        // ---- stage A: one thread per sample: tap box and in-tile weights
        if (threadIdx.x < nb) {
            const int j = threadIdx.x;
            const int64_t s = slist[eb + j];
            const float* rec = taprec + s * g.rec;
            const int* ri = reinterpret_cast<const int*>(rec);
            int lo[3], cn[3], l0[3];
            for (int a = 0; a < 3; ++a) {
                const int W = g.width[a], G = g.os[a];
                int rel0 = ri[a] - o[a];
                if (rel0 < 0) rel0 += G;
                if (g.ntiles[a] == 1) {
                    // one tile spans the axis: every tap is inside, local
                    // coordinates wrap modulo G in stage C
                    lo[a] = 0;
                    l0[a] = rel0;
                    cn[a] = W;
                } else if (rel0 < e[a]) {
                    lo[a] = 0;
                    l0[a] = rel0;
                    cn[a] = min(W, e[a] - rel0);
                } else {
                    int d = G - rel0;  // taps until the tile start
                    lo[a] = d;
                    l0[a] = 0;
                    cn[a] = d < W ? min(W - d, e[a]) : 0;
                }
                for (int k = 0; k < cn[a]; ++k) sw[(j * 3 + a) * PF_WMAX + k] = rec[g.woff[a] + lo[a] + k];
            }
            int* bx = sbox + j * 8;
            bx[0] = cn[0];
            bx[1] = cn[1];
            bx[2] = cn[2];
            bx[3] = l0[0];
            bx[4] = l0[1];
            bx[5] = l0[2];
            bx[6] = ri[3];          // frame
            bx[7] = (int)s;
        }
        __syncthreads();
        // ---- stage B: values (times conj coefficient)
        for (int q = threadIdx.x; q < nb * vcnt; q += TB) {
            const int j = q / vcnt, v = q % vcnt;
            const int u = vbase + v;
            const int cb = u / RC, rv = u % RC;
            const int r = rv / B.Cv, cv = rv % B.Cv;
            const int64_t s = sbox[j * 8 + 7];
            float2 val = ys[s * B.C + cb * B.Cv + cv];
            if (B.basis || B.b0t) {
                int64_t tread = B.b0t ? ((int64_t)perm[s] % g.npts) % B.nt : 0;
                val = cmulc(val, sample_coef(B, r, sbox[j * 8 + 6], tread));
            }
            sy[j * VC + v] = val;
        }
        __syncthreads();
        // ---- stage C: per-wave accumulation in sample order
        if (vwn > 0) {
            for (int j = 0; j < nb; ++j) {
                const int* bx = sbox + j * 8;
                const int cz = bx[0], cy = bx[1], cx = bx[2];
                const int n = cz * cy * cx;
                if (n == 0) continue;
                const int mcx = (65536 + cx - 1) / cx, mcy = (65536 + cy - 1) / cy;
                const int lz0 = bx[3], ly0 = bx[4], lx0 = bx[5];
                const float* wz = sw + (j * 3 + 0) * PF_WMAX;
                const float* wy = sw + (j * 3 + 1) * PF_WMAX;
                const float* wx = sw + (j * 3 + 2) * PF_WMAX;
                const float2* yj = sy + j * VC + vw0;
                const int npairs = n * vwn;
                for (int q = lane; q < npairs; q += 64) {
                    const int t = fdiv(q, mvw);
                    const int v = q - t * vwn;
                    const int t2 = fdiv(t, mcx);
                    const int ix = t - t2 * cx;
                    const int iz = fdiv(t2, mcy);
                    const int iy = t2 - iz * cy;
                    const float w = wz[iz] * wy[iy] * wx[ix];
                    // wrap only happens on single-tile axes (e = G there)
                    int lz = lz0 + iz, ly = ly0 + iy, lx = lx0 + ix;
                    lz -= lz >= e[0] ? e[0] : 0;
                    ly -= ly >= e[1] ? e[1] : 0;
                    lx -= lx >= e[2] ? e[2] : 0;
                    const int pt = (lz * Ty + ly) * Tx + lx;
                    const float2 val = yj[v];
                    float* d = acc + pt * PS + 2 * (vw0 + v);
                    atomicAdd(d, w * val.x);
                    atomicAdd(d + 1, w * val.y);
                }
            }
        }
    }
    __syncthreads();
    if (it.w < 0) {
        for (int i = threadIdx.x; i < tpts * vcnt; i += TB) {
            const int v = i % vcnt;
            const int pt = i / vcnt;
            const int lx = pt % Tx, ly = (pt / Tx) % Ty, lz = pt / (Tx * Ty);
            if (lx >= e[2] || ly >= e[1] || lz >= e[0]) continue;
            const int u = vbase + v;
            const int cb = u / RC, rv = u % RC;
            const int64_t gpt = ((int64_t)(o[0] + lz) * g.os[1] + (o[1] + ly)) * g.os[2] + (o[2] + lx);
            const int64_t slab = (int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + cb;
            const float* src = acc + pt * PS + 2 * v;
            grid[(slab * g.gtot + gpt) * RC + rv] = make_float2(src[0], src[1]);
        }
    } else {
        float2* dst = partial + ((int64_t)it.w * nvch + vch) * (int64_t)tpts * VC;
        for (int i = threadIdx.x; i < tpts * VC; i += TB) {
            const int v = i % VC, pt = i / VC;
            const float* src = acc + pt * PS + 2 * v;
            dst[i] = make_float2(src[0], src[1]);
        }
    }
}

// Sum the partial tiles of a heavy tile in slot order and write it out.
__global__ void __launch_bounds__(TB) k_spread_reduce(GeoParams g, BatchParams B, int VC, int nsub,
                                                      const int4* __restrict__ ritems,
                                                      const float2* __restrict__ partial,
                                                      float2* __restrict__ grid) {
    // block = (split tile, value chunk, sub-range of TB tile values); each
    // thread sums one value's partials in chunk order
    const int RC = B.R * B.Cv;
    const int V = B.Cb * RC;
    const int nvch = (V + VC - 1) / VC;
    const int iv = blockIdx.x / nsub;
    const int4 it = ritems[iv / nvch];
    const int vch = iv % nvch;
    const int vbase = vch * VC;
    const int vcnt = min(VC, V - vbase);
    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tx = (int)(tile % g.ntiles[2]);
    const int ty = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
    const int ox = tx * Tx, oy = ty * Ty, oz = tz * Tz;
    const int ex = min(Tx, g.os[2] - ox), ey = min(Ty, g.os[1] - oy), ez = min(Tz, g.os[0] - oz);
    const int tpts = Tx * Ty * Tz;
    {
        const int i = (blockIdx.x % nsub) * TB + threadIdx.x;
        if (i >= tpts * vcnt) return;
        const int v = i % vcnt;
        const int pt = i / vcnt;
        const int lx = pt % Tx, ly = (pt / Tx) % Ty, lz = pt / (Tx * Ty);
        if (lx >= ex || ly >= ey || lz >= ez) return;
        float2 a = make_float2(0.f, 0.f);
#pragma unroll 4
        for (int sl = 0; sl < it.z; ++sl) {
            const float2* src = partial + ((int64_t)(it.y + sl) * nvch + vch) * (int64_t)tpts * VC;
            const float2 x = src[(int64_t)pt * VC + v];
            a.x += x.x;
            a.y += x.y;
        }
        const int u = vbase + v;
        const int cb = u / RC, rv = u % RC;
        const int64_t gpt = ((int64_t)(oz + lz) * g.os[1] + (oy + ly)) * g.os[2] + (ox + lx);
        const int64_t slab = (int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + cb;
        grid[(slab * g.gtot + gpt) * RC + rv] = a;
    }
}

// -------------------------------------------------------------------------
// Register-resident spread for batches of >= 16 values per grid point.
// One wave owns one tile (TZ x TY x 2*TXL points of the plan's tiling) for
// a chunk of 32 values of the flat per-point vector (Cb, R, Cv): lane =
// (value, x half), each lane keeps its TZ*TY*TXL complex accumulators in
// VGPRs and walks the tile's sample list in order.
//
// The list is consumed in batches of SPREAD_EB entries: first every lane
// fetches one entry (sample index, tap offsets relative to the tile, tap
// weights) into LDS, then the lanes fetch the batch's sample values (all
// independent loads, issued back to back), then the wave accumulates from
// LDS.  One global round trip per batch instead of two dependent ones per
// entry.  Tap rows of an entry are uniform across the wave (scalar branches
// skip rows outside the footprint); each useful tap costs two FMAs.  No
// atomics: the summation order is the sample order, so results are
// run-to-run bit-identical.
// -------------------------------------------------------------------------
constexpr int SPREAD_VL = 32;
constexpr int SPREAD_EB = 32;  // 11.7 KB of LDS per wave: 12 waves per CU fit

#ifndef PF_SPREAD_UNROLL
#define PF_SPREAD_UNROLL 1
#endif
#ifndef PF_SPREAD_REG_MIN
#define PF_SPREAD_REG_MIN 1  // below: the LDS-staged spread (measured slower from 1 value per point up)
#endif
#ifndef PF_SPREAD_PIPE
#define PF_SPREAD_PIPE 1  // plain 32-coil batches: software-pipelined copies
#endif
#ifndef PF_SPREAD_MFMA
#define PF_SPREAD_MFMA 1  // plain 32-coil batches on 128-point 3-D tiles: matrix-core accumulation
#endif
#ifndef PF_SPREAD_EXPAND
#define PF_SPREAD_EXPAND 1  // fused adjoint with B0 / basis: the sort expands the R*C values per sample
#endif
#ifndef PF_SPREAD_WREG
#define PF_SPREAD_WREG 1  // an entry's z/y weights in registers (measured faster)
#endif

// 3-D tiles (few z/y rows) and the 8-row 2-D tiles keep an entry's z/y
// weights in registers and are compiled for >= 3 waves per SIMD (168 VGPRs;
// the 32-value 8 x 16 instance spills 3-5 and still measured 3-10 us faster
// on a 64-value 2-D spread than reading the weights from LDS per row); the
// other 2-D tiles (longer y columns or x rows) and 1-D read them from LDS.
template <int ZY>  // 100 * TZ + TY
constexpr bool spread_wreg() {
    return PF_SPREAD_WREG && (ZY / 100 > 1 || ZY == 108) && ZY / 100 + ZY % 100 <= 12;
}

// VL values per wave (32, 16 or 8 for small batches): lane = (value, part),
// 64 / VL parts; a part is (z part, x part): ZP z parts of TZ / ZP rows (ZP > 1
// only where the tile's x extent is narrower than 64 / VL) and the rest x
// parts of TXL points each.
template <int TZ, int TY, int TXL, int VL = SPREAD_VL, int ZP = 1>
__global__ void __launch_bounds__(64, spread_wreg<100 * TZ + TY>() ? 3 : 1) k_spread_reg(GeoParams g, BatchParams B, const int4* __restrict__ items,
                                                   const int32_t* __restrict__ slist,
                                                   const float* __restrict__ taprec,
                                                   const int32_t* __restrict__ perm,
                                                   const float2* __restrict__ ys, float2* __restrict__ grid,
                                                   float2* __restrict__ partial) {
    static_assert(VL == 8 || VL == 16 || VL == 32, "values per wave");
    static_assert((64 / VL) % ZP == 0 && TZ % ZP == 0, "z parts");
    constexpr int P = 64 / VL;    // parts per value
    constexpr int XP = P / ZP;    // x parts
    constexpr int TZL = TZ / ZP;  // z rows per lane
    constexpr int EB = SPREAD_EB;
    constexpr int TX = XP * TXL;
    constexpr int NW = TZ + TY + TX;  // tile-local tap weights of one entry
    static_assert(TZ <= 32 && TY <= 32, "row masks hold 32 bits per axis");
    __shared__ int s_idx[EB];
    __shared__ int s_frame[EB];
    __shared__ unsigned s_zmask[EB];           // tile z rows inside the footprint
    __shared__ unsigned s_ymask[EB];           // tile y rows inside the footprint
    __shared__ float s_w[EB][NW];              // weights at the tile's z, y, x positions (0 outside)
    __shared__ __attribute__((aligned(16))) float2 s_val[EB][VL];  // values times conj(coefficient)

    const int RC = B.R * B.Cv;
    const int V = B.Cb * RC;
    const int nvch = (V + VL - 1) / VL;
    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int4 it = items[bid / nvch];
    const int vch = bid % nvch;
    const int lane = threadIdx.x;
    const int v = lane & (VL - 1);
    const int part = lane / VL;
    const int half = part % XP;  // x part
    const int zp = part / XP;    // z part
    const int u = vch * VL + v;
    const bool vok = u < V;
    const int uu = vok ? u : V - 1;
    const int cb = uu / RC, rv = uu % RC;
    const int r = rv / B.Cv, cv = rv % B.Cv;
    const int ycol = cb * B.Cv + cv;
    const bool dma = VL == SPREAD_VL && B.dma != 0;  // whole 32-coil rows only

    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tix = (int)(tile % g.ntiles[2]);
    const int tiy = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int oz = tiz * g.tile[0], oy = tiy * g.tile[1], ox = tix * g.tile[2];
    const int ez = min(g.tile[0], g.os[0] - oz), ey = min(g.tile[1], g.os[1] - oy);
    const int ex = min(g.tile[2], g.os[2] - ox);

    float2 acc[TZL][TY][TXL];
#pragma unroll
    for (int a = 0; a < TZL; ++a)
#pragma unroll
        for (int b = 0; b < TY; ++b)
#pragma unroll
            for (int k = 0; k < TXL; ++k) acc[a][b][k] = make_float2(0.f, 0.f);

    for (int e0 = it.y; e0 < it.z; e0 += EB) {
        const int n = min(EB, it.z - e0);
        // phase 1: one entry per lane -> its weights at this tile's positions
        if (lane < n) {
            const int s = slist[e0 + lane];
            const float* rec = taprec + (int64_t)s * g.rec;
            const int* ri = reinterpret_cast<const int*>(rec);
            // tap index of tile position p along an axis: (o + p - start) mod G
            int bz = oz - ri[0];
            bz += bz < 0 ? g.os[0] : 0;
            int by = oy - ri[1];
            by += by < 0 ? g.os[1] : 0;
            int bx = ox - ri[2];
            bx += bx < 0 ? g.os[2] : 0;
            unsigned mz = 0, my = 0;
#pragma unroll
            for (int a = 0; a < TZ; ++a) {
                int i = bz + a;
                i -= i >= g.os[0] ? g.os[0] : 0;
                const bool in = i < g.width[0];
                s_w[lane][a] = in ? rec[g.woff[0] + i] : 0.f;
                mz |= in ? 1u << a : 0u;
            }
#pragma unroll
            for (int b = 0; b < TY; ++b) {
                int i = by + b;
                i -= i >= g.os[1] ? g.os[1] : 0;
                const bool in = i < g.width[1];
                s_w[lane][TZ + b] = in ? rec[g.woff[1] + i] : 0.f;
                my |= in ? 1u << b : 0u;
            }
#pragma unroll
            for (int x = 0; x < TX; ++x) {
                int i = bx + x;
                i -= i >= g.os[2] ? g.os[2] : 0;
                s_w[lane][TZ + TY + x] = i < g.width[2] ? rec[g.woff[2] + i] : 0.f;
            }
            s_idx[lane] = s;
            s_frame[lane] = ri[3];
            s_zmask[lane] = mz;
            s_ymask[lane] = my;
        }
        __syncthreads();
        // phase 2: the batch's values
        if (VL == SPREAD_VL && dma) {
            // plain values, 32 whole coils per wave: copied global -> LDS
            // without registers, 16 lanes x 16 bytes per entry, 4 entries per
            // instruction, all 8 instructions in flight at once
            const int sub = lane >> 4, p16 = lane & 15;
            const char* ybase = reinterpret_cast<const char*>(ys);
            // not unrolled: one address live at a time (the accumulators hold
            // the register budget); the copies still overlap, nothing waits
            // on their data until the barrier
#pragma unroll 1
            for (int q = 0; q < EB / 4; ++q) {
                const int j = q * 4 + sub;
                const unsigned sj = (unsigned)s_idx[j < n ? j : 0];  // slots past n are never read
                // 32-bit byte offset (the launcher checks the samples fit in 4 GiB)
                const unsigned off = (sj * (unsigned)B.C + (unsigned)(vch * VL)) * 8u + (unsigned)p16 * 16u;
                __builtin_amdgcn_global_load_lds(ybase + off, (__attribute__((address_space(3))) void*)&s_val[q * 4][0],
                                                 16, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(0);
        } else {
            // lane (value, part) takes entries part, part + P, ...
            for (int j = part; j < n; j += P) {
                const int s = s_idx[j];
                float2 val = vok ? ys[(int64_t)s * B.C + ycol] : make_float2(0.f, 0.f);
                if (B.basis || B.b0t) {
                    const int64_t tread = B.b0t ? ((int64_t)perm[s] % g.npts) % B.nt : 0;
                    val = cmulc(val, sample_coef(B, r, s_frame[j], tread));
                }
                s_val[j][v] = val;
            }
        }
        __syncthreads();
        // phase 3: accumulate; rows outside the footprint are skipped by
        // scalar tests of the entry's row masks
#pragma unroll PF_SPREAD_UNROLL
        for (int j = 0; j < n; ++j) {
            const unsigned mz = __builtin_amdgcn_readfirstlane(s_zmask[j]);
            const unsigned my = __builtin_amdgcn_readfirstlane(s_ymask[j]);
            const float2 val = s_val[j][v];
            const float* w = s_w[j];
            // the value times this lane's x weights, once per entry: a row
            // pair then costs one z*y product and two FMAs per x point (wide
            // x runs keep the weights instead, for registers)
            constexpr bool VX = TXL <= 4;
            float2 vxk[TXL];
            float wxl[TXL];
#pragma unroll
            for (int k = 0; k < TXL; ++k) {
                const float wx = w[TZ + TY + half * TXL + k];
                if constexpr (VX) vxk[k] = make_float2(wx * val.x, wx * val.y);
                else wxl[k] = wx;
            }
            // all z/y weights of the entry up front (vector LDS reads), so the
            // row loop does not wait on LDS
            constexpr bool WR = spread_wreg<100 * TZ + TY>();
            float wzr[WR ? TZ : 1], wyr[WR ? TY : 1];
            if constexpr (WR) {
#pragma unroll
                for (int a = 0; a < TZ; ++a) wzr[a] = w[a];
#pragma unroll
                for (int b = 0; b < TY; ++b) wyr[b] = w[TZ + b];
            }
#pragma unroll
            for (int a = 0; a < TZL; ++a) {
                // the row pass runs when any z part's row a is inside the
                // footprint; a lane whose own row is outside reads weight 0
                unsigned rows = 0;
#pragma unroll
                for (int zq = 0; zq < ZP; ++zq) rows |= mz >> (zq * TZL + a);
                if (!(rows & 1u)) continue;
                float wz;
                if constexpr (WR) {
                    wz = wzr[a];
#pragma unroll
                    for (int zq = 1; zq < ZP; ++zq) wz = zp == zq ? wzr[WR ? zq * TZL + a : 0] : wz;
                } else {
                    wz = w[zp * TZL + a];
                }
#pragma unroll
                for (int b = 0; b < TY; ++b) {
                    if (!((my >> b) & 1u)) continue;
                    const float wzy = wz * (WR ? wyr[WR ? b : 0] : w[TZ + b]);
                    if constexpr (VX) {
#pragma unroll
                        for (int k = 0; k < TXL; ++k) {
                            acc[a][b][k].x += wzy * vxk[k].x;
                            acc[a][b][k].y += wzy * vxk[k].y;
                        }
                    } else {
                        const float vx = wzy * val.x, vy = wzy * val.y;
#pragma unroll
                        for (int k = 0; k < TXL; ++k) {
                            acc[a][b][k].x += wxl[k] * vx;
                            acc[a][b][k].y += wxl[k] * vy;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

    if (!vok) return;
    if (it.w < 0) {
        const int64_t slab = (int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + cb;
        float2* base = grid + slab * g.gtot * RC + rv;
#pragma unroll
        for (int a = 0; a < TZL; ++a)
#pragma unroll
            for (int b = 0; b < TY; ++b)
#pragma unroll
                for (int k = 0; k < TXL; ++k) {
                    const int lx = half * TXL + k;
                    const int lz = zp * TZL + a;
                    if (lz < ez && b < ey && lx < ex) {
                        const int64_t gpt = ((int64_t)(oz + lz) * g.os[1] + (oy + b)) * g.os[2] + (ox + lx);
                        base[gpt * RC] = acc[a][b][k];
                    }
                }
    } else {
        const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
        float2* dst = partial + ((int64_t)it.w * nvch + vch) * (int64_t)(Tx * Ty * Tz) * VL + v;
#pragma unroll
        for (int a = 0; a < TZL; ++a)
#pragma unroll
            for (int b = 0; b < TY; ++b)
#pragma unroll
                for (int k = 0; k < TXL; ++k) {
                    const int lx = half * TXL + k;
                    const int lz = zp * TZL + a;
                    if (lz < Tz && b < Ty && lx < Tx) dst[(int64_t)((lz * Ty + b) * Tx + lx) * VL] = acc[a][b][k];
                }
    }
}

// -------------------------------------------------------------------------
// Software-pipelined register spread for plain values in whole 32-coil
// chunks (no basis, no B0).  Same tile ownership, lane roles and summation
// order as k_spread_reg; the difference is where the wave waits.  Batches of
// SPIPE_EB entries alternate between two LDS buffer pairs: while the wave
// accumulates batch b from one pair, the tap records and the values of
// batch b + 1 are already being copied global -> LDS into the other pair
// (global_load_lds, no registers held), and the sample ids of batch b + 2 are
// in flight.  The wave then waits once per batch, for copies issued a whole
// batch of arithmetic earlier.
// -------------------------------------------------------------------------
constexpr int SPIPE_EB = 16;
constexpr int SPIPE_RP = 8;  // 16-byte record pieces per entry slot (a record has at most 4 + 3 * 8 dwords)
static_assert(4 + 3 * PF_WMAX <= 4 * SPIPE_RP, "tap record exceeds its LDS slot");

// an entry's row: its TZ + TY + TX weights at the tile's positions, then the
// bit masks of the tile z and y rows its footprint covers; padded to 16 bytes
template <int TZ, int TY, int TXL>
constexpr int spipe_row() { return (TZ + TY + 2 * TXL + 2 + 3) / 4 * 4; }

// copies of one batch: each entry's tap record (rec4 pieces of 16 bytes; the
// slot's unused pieces take a harmless copy of piece 0) and its 32 values;
// sid holds lane j's sample id for entries j < n
__device__ __forceinline__ void spipe_issue(float4 (&s_rec)[SPIPE_EB][SPIPE_RP], float2 (&val)[SPIPE_EB][SPREAD_VL],
                                            int sid, int n, int rec4, const float* __restrict__ taprec,
                                            const char* __restrict__ ybase, int C, int vch) {
    const int lane = threadIdx.x;
    const int c = lane & (SPIPE_RP - 1);
    const int cs = c < rec4 ? c : 0;
#pragma unroll
    for (int k = 0; k < SPIPE_EB * SPIPE_RP / 64; ++k) {
        const int e = k * (64 / SPIPE_RP) + (lane >> 3);
        const int se = __shfl(sid, e < n ? e : 0);
        const float* src = taprec + (int64_t)se * (rec4 * 4) + 4 * cs;
        if (k * (64 / SPIPE_RP) < n)
            __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)&s_rec[k * (64 / SPIPE_RP)][0],
                                             16, 0, 0);
    }
    // values: 16 lanes x 16 bytes per entry, 4 entries per copy
    const int sub = lane >> 4, p16 = lane & 15;
#pragma unroll
    for (int q = 0; q < SPIPE_EB / 4; ++q) {
        const int j = q * 4 + sub;
        const unsigned sj = (unsigned)__shfl(sid, j < n ? j : 0);
        const unsigned off = (sj * (unsigned)C + (unsigned)(vch * SPREAD_VL)) * 8u + (unsigned)p16 * 16u;
        if (q * 4 < n)
            __builtin_amdgcn_global_load_lds(ybase + off, (__attribute__((address_space(3))) void*)&val[q * 4][0], 16,
                                             0, 0);
    }
}

// one entry per lane: its record -> its weights at this tile's positions
// (0 outside the footprint) and the tile rows its footprint covers
template <int TZ, int TY, int TXL>
__device__ __forceinline__ void spipe_weights(const float4 (&s_rec)[SPIPE_EB][SPIPE_RP], int n, const GeoParams& g,
                                              int oz, int oy, int ox,
                                              float (&s_w)[SPIPE_EB][spipe_row<TZ, TY, TXL>()]) {
    constexpr int TX = 2 * TXL;
    const int lane = threadIdx.x;
    if (lane >= n) return;
    const float* rec_l = reinterpret_cast<const float*>(&s_rec[lane][0]);
    const int* ri = reinterpret_cast<const int*>(rec_l);
    int bz = oz - ri[0];
    bz += bz < 0 ? g.os[0] : 0;
    int by = oy - ri[1];
    by += by < 0 ? g.os[1] : 0;
    int bx = ox - ri[2];
    bx += bx < 0 ? g.os[2] : 0;
    unsigned mz = 0, my = 0;
#pragma unroll
    for (int a = 0; a < TZ; ++a) {
        int i = bz + a;
        i -= i >= g.os[0] ? g.os[0] : 0;
        const bool in = i < g.width[0];
        s_w[lane][a] = in ? rec_l[g.woff[0] + i] : 0.f;
        mz |= in ? 1u << a : 0u;
    }
#pragma unroll
    for (int b = 0; b < TY; ++b) {
        int i = by + b;
        i -= i >= g.os[1] ? g.os[1] : 0;
        const bool in = i < g.width[1];
        s_w[lane][TZ + b] = in ? rec_l[g.woff[1] + i] : 0.f;
        my |= in ? 1u << b : 0u;
    }
#pragma unroll
    for (int x = 0; x < TX; ++x) {
        int i = bx + x;
        i -= i >= g.os[2] ? g.os[2] : 0;
        s_w[lane][TZ + TY + x] = i < g.width[2] ? rec_l[g.woff[2] + i] : 0.f;
    }
    constexpr int NW = TZ + TY + TX;
    s_w[lane][NW] = __uint_as_float(mz);
    s_w[lane][NW + 1] = __uint_as_float(my);
}

// the batch's n entries into the lane's accumulators, in entry order
template <int TZ, int TY, int TXL>
__device__ __forceinline__ void spipe_accumulate(const float2 (&val_b)[SPIPE_EB][SPREAD_VL], int n,
                                                 const float (&s_w)[SPIPE_EB][spipe_row<TZ, TY, TXL>()],
                                                 float2 (&acc)[TZ][TY][TXL]) {
    const int lane = threadIdx.x;
    const int v = lane & (SPREAD_VL - 1);
    const int half = lane / SPREAD_VL;
    constexpr bool WR = spread_wreg<100 * TZ + TY>();
    constexpr bool VX = TXL <= 4;
    for (int j = 0; j < n; ++j) {
        constexpr int NW = TZ + TY + 2 * TXL;
        const unsigned mz = __builtin_amdgcn_readfirstlane(__float_as_uint(s_w[j][NW]));
        const unsigned my = __builtin_amdgcn_readfirstlane(__float_as_uint(s_w[j][NW + 1]));
        const float2 val = val_b[j][v];
        const float* w = s_w[j];
        float2 vxk[TXL];
        float wxl[TXL];
#pragma unroll
        for (int k = 0; k < TXL; ++k) {
            const float wx = w[TZ + TY + half * TXL + k];
            if constexpr (VX) vxk[k] = make_float2(wx * val.x, wx * val.y);
            else wxl[k] = wx;
        }
        float wzr[WR ? TZ : 1], wyr[WR ? TY : 1];
        if constexpr (WR) {
#pragma unroll
            for (int a = 0; a < TZ; ++a) wzr[a] = w[a];
#pragma unroll
            for (int b = 0; b < TY; ++b) wyr[b] = w[TZ + b];
        }
#pragma unroll
        for (int a = 0; a < TZ; ++a) {
            if (!((mz >> a) & 1u)) continue;
            const float wz = WR ? wzr[WR ? a : 0] : w[a];
#pragma unroll
            for (int b = 0; b < TY; ++b) {
                if (!((my >> b) & 1u)) continue;
                const float wzy = wz * (WR ? wyr[WR ? b : 0] : w[TZ + b]);
                if constexpr (VX) {
#pragma unroll
                    for (int k = 0; k < TXL; ++k) {
                        acc[a][b][k].x += wzy * vxk[k].x;
                        acc[a][b][k].y += wzy * vxk[k].y;
                    }
                } else {
                    const float vx = wzy * val.x, vy = wzy * val.y;
#pragma unroll
                    for (int k = 0; k < TXL; ++k) {
                        acc[a][b][k].x += wxl[k] * vx;
                        acc[a][b][k].y += wxl[k] * vy;
                    }
                }
            }
        }
    }
}

template <int TZ, int TY, int TXL>
__global__ void __launch_bounds__(64, spread_wreg<100 * TZ + TY>() ? 3 : 1)
    k_spread_pipe(GeoParams g, BatchParams B, const int4* __restrict__ items, const int32_t* __restrict__ slist,
                  const float* __restrict__ taprec, const float2* __restrict__ ys, float2* __restrict__ grid,
                  float2* __restrict__ partial) {
    constexpr int VL = SPREAD_VL;
    constexpr int EB = SPIPE_EB;
    __shared__ __attribute__((aligned(16))) float4 s_rec[EB][SPIPE_RP];
    __shared__ __attribute__((aligned(16))) float2 s_val0[EB][VL];
    __shared__ __attribute__((aligned(16))) float2 s_val1[EB][VL];
    __shared__ __attribute__((aligned(16))) float s_w[EB][spipe_row<TZ, TY, TXL>()];

    const int nvch = B.C / VL;  // plain values: one per coil
    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int4 it = items[bid / nvch];
    const int vch = bid % nvch;
    const int lane = threadIdx.x;
    const int v = lane & (VL - 1);
    const int half = lane / VL;
    const int u = vch * VL + v;

    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tix = (int)(tile % g.ntiles[2]);
    const int tiy = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int oz = tiz * g.tile[0], oy = tiy * g.tile[1], ox = tix * g.tile[2];
    const int ez = min(g.tile[0], g.os[0] - oz), ey = min(g.tile[1], g.os[1] - oy);
    const int ex = min(g.tile[2], g.os[2] - ox);
    const int rec4 = g.rec / 4;
    const char* ybase = reinterpret_cast<const char*>(ys);

    float2 acc[TZ][TY][TXL];
#pragma unroll
    for (int a = 0; a < TZ; ++a)
#pragma unroll
        for (int b = 0; b < TY; ++b)
#pragma unroll
            for (int k = 0; k < TXL; ++k) acc[a][b][k] = make_float2(0.f, 0.f);

    // batch i covers entries [it.y + i*EB, ...): its records land in s_rec,
    // its values in s_val0 (even i) or s_val1 (odd i).  Per batch: wait for
    // its copies, turn records into weights, start the next batch's copies
    // (records overwrite s_rec, which is no longer read), accumulate.
    const int cnt = it.z - it.y;
    const int32_t* sl = slist + it.y;  // the item's entries
    int sidA = lane < cnt ? sl[lane] : 0;
    int sidB = lane < cnt - EB ? sl[EB + lane] : 0;
    spipe_issue(s_rec, s_val0, sidA, min(EB, cnt), rec4, taprec, ybase, B.C, vch);
    for (int e0 = 0; e0 < cnt; e0 += 2 * EB) {
        {
            const int n = min(EB, cnt - e0);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            spipe_weights<TZ, TY, TXL>(s_rec, n, g, oz, oy, ox, s_w);
            __syncthreads();
            const int n1 = min(EB, cnt - e0 - EB);
            if (n1 > 0) {
                // ids first: a reload of a spilled register before them waits on
                // nothing but itself
                const int sidC = lane < cnt - e0 - 2 * EB ? sl[e0 + 2 * EB + lane] : 0;
                spipe_issue(s_rec, s_val1, sidB, n1, rec4, taprec, ybase, B.C, vch);
                sidA = sidC;
            }
            spipe_accumulate<TZ, TY, TXL>(s_val0, n, s_w, acc);
        }
        if (e0 + EB >= cnt) break;
        {
            const int n = min(EB, cnt - e0 - EB);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            spipe_weights<TZ, TY, TXL>(s_rec, n, g, oz, oy, ox, s_w);
            __syncthreads();
            const int n2 = min(EB, cnt - e0 - 2 * EB);
            if (n2 > 0) {
                const int sidC = lane < cnt - e0 - 3 * EB ? sl[e0 + 3 * EB + lane] : 0;
                spipe_issue(s_rec, s_val0, sidA, n2, rec4, taprec, ybase, B.C, vch);
                sidB = sidC;
            }
            spipe_accumulate<TZ, TY, TXL>(s_val1, n, s_w, acc);
        }
    }

    if (it.w < 0) {
        const int64_t slab = (int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + u / B.Cv;
        float2* base = grid + slab * g.gtot * B.Cv + u % B.Cv;
#pragma unroll
        for (int a = 0; a < TZ; ++a)
#pragma unroll
            for (int b = 0; b < TY; ++b)
#pragma unroll
                for (int k = 0; k < TXL; ++k) {
                    const int lx = half * TXL + k;
                    if (a < ez && b < ey && lx < ex) {
                        const int64_t gpt = ((int64_t)(oz + a) * g.os[1] + (oy + b)) * g.os[2] + (ox + lx);
                        base[gpt * B.Cv] = acc[a][b][k];
                    }
                }
    } else {
        const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
        float2* dst = partial + ((int64_t)it.w * nvch + vch) * (int64_t)(Tx * Ty * Tz) * VL + v;
#pragma unroll
        for (int a = 0; a < TZ; ++a)
#pragma unroll
            for (int b = 0; b < TY; ++b)
#pragma unroll
                for (int k = 0; k < TXL; ++k) {
                    const int lx = half * TXL + k;
                    if (a < Tz && b < Ty && lx < Tx) dst[(int64_t)((a * Ty + b) * Tx + lx) * VL] = acc[a][b][k];
                }
    }
}

// -------------------------------------------------------------------------
// Matrix-core spread: the same items, entry batches, copies and summation
// order over batches as k_spread_pipe, with the accumulation of a batch done
// by MFMA.  For a batch of SPIPE_EB = 16 entries the tile update is a product
//
//     acc[p, n] += sum_k A[p, k] * B[k, n],   A[p, k] = wz * wy * wx of entry k
//                                             at tile point p (0 outside its
//                                             footprint), B[k, n] = re/im part
//                                             n of entry k's 32 values,
//
// p over the tile's 128 points (4 row blocks of 32), n over 64 real columns
// (2 blocks of 32): 8 v_mfma_f32_32x32x16_f16 per product.  FP32 operands are
// split into fp16 pairs, x = hi + lo with |x - hi - lo| <= 2^-22 |x|, and
// three products hi*hi + hi*lo + lo*hi accumulate in FP32 (the dropped lo*lo
// is below 2^-22 relative), so the result matches FP32 arithmetic to rounding
// level while the tile's arithmetic runs on the matrix cores instead of the
// vector ALUs.  fp16 range: weights are normalised by the kernel's peak
// value (w <= 1), values by a power of two per work item and value chunk,
// taken from the largest |value| among its samples (k_tile_max over the
// per-sample maxima the sort or k_smax wrote) so that scaled values stay
// below 2^13; both are undone when the tile is written.  Deterministic: a fixed instruction sequence per tile.
// -------------------------------------------------------------------------
// per sorted sample t and 32-value chunk v: max |re|, |im| over the chunk's
// values, smax[t * nvch + v]; 16 lanes per (t, v), one 16-byte piece each
__global__ void __launch_bounds__(TB) k_smax(int64_t mtot, int C, const float2* __restrict__ ys,
                                             float* __restrict__ smax) {
    const int nvch = C / SPREAD_VL;
    const int64_t gid = (int64_t)blockIdx.x * TB + threadIdx.x;
    const int64_t row = gid >> 4;  // (t, v) pair
    const int piece = (int)(gid & 15);
    float m = 0.f;
    if (row < mtot * nvch) {
        const int64_t t = row / nvch;
        const int v = (int)(row - t * nvch);
        const float4 q = reinterpret_cast<const float4*>(ys + t * C + (int64_t)v * SPREAD_VL)[piece];
        m = fmaxf(fmaxf(fabsf(q.x), fabsf(q.y)), fmaxf(fabsf(q.z), fabsf(q.w)));
    }
    int x = __float_as_int(m);
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    if (piece == 15 && row < mtot * nvch) smax[row] = __int_as_float(x);
}

typedef _Float16 pf_half2 __attribute__((ext_vector_type(2)));
typedef _Float16 pf_half8 __attribute__((ext_vector_type(8)));
typedef float pf_f2 __attribute__((ext_vector_type(2)));
typedef float pf_f16 __attribute__((ext_vector_type(16)));

// (a, b) -> fp16 pairs hi = rn(x), lo = rn(x - hi)
__device__ __forceinline__ void smfma_split(float a, float b, pf_half2& hi, pf_half2& lo) {
    const pf_f2 x = {a, b};
    hi = __builtin_convertvector(x, pf_half2);
    lo = __builtin_convertvector(x - __builtin_convertvector(hi, pf_f2), pf_half2);
}

__device__ __forceinline__ pf_half8 smfma_cat(pf_half2 a, pf_half2 b, pf_half2 c, pf_half2 d) {
    return __builtin_shufflevector(__builtin_shufflevector(a, b, 0, 1, 2, 3), __builtin_shufflevector(c, d, 0, 1, 2, 3),
                                   0, 1, 2, 3, 4, 5, 6, 7);
}

#ifndef PF_SMFMA_OCC
#define PF_SMFMA_OCC 2  // waves per SIMD the register budget is sized for
#endif
constexpr int SMFMA_NCOPY = SPIPE_EB * SPIPE_RP / 64 + SPIPE_EB / 4 + 1;  // copy instructions per batch

// max over the wave of non-negative floats (integer order of their bits),
// by DPP row shifts and row broadcasts; the result is read from lane 63
__device__ __forceinline__ float smfma_wave_max(float v) {
    int x = __float_as_int(v);
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

// per spread work item and 32-value chunk (index item * nvch + chunk, the
// spread's block index): the largest of its samples' smax, one wave each
__global__ void __launch_bounds__(TB) k_tile_max(int nitems, int nvch, const int4* __restrict__ items,
                                                 const int32_t* __restrict__ slist, const float* __restrict__ smax,
                                                 float* __restrict__ tmax) {
    const int64_t wid = ((int64_t)blockIdx.x * TB + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wid >= (int64_t)nitems * nvch) return;  // whole waves
    const int4 it = items[wid / nvch];
    const int vch = (int)(wid % nvch);
    float m = 0.f;
    for (int e = it.y + lane; e < it.z; e += 64) m = fmaxf(m, smax[(int64_t)slist[e] * nvch + vch]);
    m = smfma_wave_max(m);
    if (lane == 0) tmax[wid] = m;
}


// copies of one batch into one record slot and one value slot, plus the 16
// sample ids of a later batch (ids_src) into an id slot: always SMFMA_NCOPY
// copy instructions (entries past n repeat entry 0), so the wait for a batch
// can count the copies issued after it; lane j < 16 of sid holds entry j's
// sample id
__device__ __forceinline__ void smfma_issue(float4 (&s_rec)[SPIPE_EB][SPIPE_RP], float2 (&val)[SPIPE_EB][SPREAD_VL],
                                            int (&s_id)[SPIPE_EB], const int32_t* __restrict__ ids_src, int sid,
                                            int n, int rec4, const float* __restrict__ taprec,
                                            const char* __restrict__ ybase, int C, int vch) {
    const int lane = threadIdx.x;
    const int c = lane & (SPIPE_RP - 1);
    const int cs = c < rec4 ? c : 0;
#pragma unroll
    for (int k = 0; k < SPIPE_EB * SPIPE_RP / 64; ++k) {
        const int e = k * (64 / SPIPE_RP) + (lane >> 3);
        const int se = __shfl(sid, e < n ? e : 0);
        const float* src = taprec + (int64_t)se * (rec4 * 4) + 4 * cs;
        __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)&s_rec[k * (64 / SPIPE_RP)][0],
                                         16, 0, 0);
    }
    const int sub = lane >> 4, p16 = lane & 15;
#pragma unroll
    for (int q = 0; q < SPIPE_EB / 4; ++q) {
        const int j = q * 4 + sub;
        const unsigned sj = (unsigned)__shfl(sid, j < n ? j : 0);
        const unsigned off = (sj * (unsigned)C + (unsigned)(vch * SPREAD_VL)) * 8u + (unsigned)p16 * 16u;
        __builtin_amdgcn_global_load_lds(ybase + off, (__attribute__((address_space(3))) void*)&val[q * 4][0], 16, 0,
                                         0);
    }
    // the ids last: the slot still held this batch's ids (sid), and sid has
    // been consumed (the shuffles above waited for its read) before the copy
    // can overwrite it
    if (lane < SPIPE_EB)
        __builtin_amdgcn_global_load_lds(ids_src + lane, (__attribute__((address_space(3))) void*)&s_id[0], 4, 0, 0);
}

// the batch's weight rows: entry e's weights at this tile's TZ z, TY y and TX
// x positions (0 outside the footprint; the x weights times wnorm), rows of
// entries n .. 15 zero.  Four lanes per entry, one 16-byte piece of the row
// each; every record read is issued before any select or store (a position
// outside the footprint reads the entry's first weight of that axis and
// drops it).
template <int TZ, int TY, int TX>
__device__ __forceinline__ void smfma_weights(const float4 (&s_rec)[SPIPE_EB][SPIPE_RP], int n, const GeoParams& g,
                                              int oz, int oy, int ox, float wnorm,
                                              float (&s_w)[SPIPE_EB][TZ + TY + TX]) {
    constexpr int NW = TZ + TY + TX;
    static_assert(NW == 16 && TZ % 4 == 0 && TY % 4 == 0 && TX % 4 == 0, "one axis per 16-byte row piece");
    const int lane = threadIdx.x;
    const int e = lane >> 2, k0 = 4 * (lane & 3);  // entry, first slot of the piece
    const int a = k0 < TZ ? 0 : (k0 < TZ + TY ? 1 : 2);
    const int p0 = k0 - (a == 0 ? 0 : (a == 1 ? TZ : TZ + TY));  // tile position of the first slot
    // selected from values held in scalar registers (a per-lane index into
    // the parameters, or a select the compiler turns into one, would read
    // the argument block through vector memory)
    int os = __builtin_amdgcn_readfirstlane(g.os[2]), W = __builtin_amdgcn_readfirstlane(g.width[2]);
    int wo = __builtin_amdgcn_readfirstlane(g.woff[2]);
    if (a == 1) {
        os = __builtin_amdgcn_readfirstlane(g.os[1]);
        W = __builtin_amdgcn_readfirstlane(g.width[1]);
        wo = __builtin_amdgcn_readfirstlane(g.woff[1]);
    }
    if (a == 0) {
        os = __builtin_amdgcn_readfirstlane(g.os[0]);
        W = __builtin_amdgcn_readfirstlane(g.width[0]);
        wo = __builtin_amdgcn_readfirstlane(g.woff[0]);
    }
    const int o = a == 0 ? oz : (a == 1 ? oy : ox);
    const float* rec_l = reinterpret_cast<const float*>(&s_rec[e][0]);
    const int* ri = reinterpret_cast<const int*>(rec_l);
    int base = o - (a == 0 ? ri[0] : (a == 1 ? ri[1] : ri[2]));
    base += base < 0 ? os : 0;
    const bool live = e < n;
    float wv[4];
    bool in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int i = base + p0 + k;
        i -= i >= os ? os : 0;
        in[k] = live && i < W;
        wv[k] = rec_l[wo + (in[k] ? i : 0)];
    }
    const float sc = a == 2 ? wnorm : 1.0f;
    reinterpret_cast<float4*>(s_w[e])[lane & 3] =
        make_float4(in[0] ? wv[0] * sc : 0.f, in[1] ? wv[1] * sc : 0.f, in[2] ? wv[2] * sc : 0.f,
                    in[3] ? wv[3] * sc : 0.f);
}

// tile TZ x TY x TX = 128 points; row block mb holds z rows mb * ZR ..
// (ZR = 32 / (TY * TX)), row i of a block is (z - mb * ZR, y, x) row-major.
// Copies run two batches ahead: records in two slots, values in three.
template <int TZ, int TY, int TX>
__global__ void __launch_bounds__(64, PF_SMFMA_OCC)
    k_spread_mfma(GeoParams g, BatchParams B, const int4* __restrict__ items, const int32_t* __restrict__ slist,
                  const float* __restrict__ taprec, const float2* __restrict__ ys, const float* __restrict__ tmax,
                  float2* __restrict__ grid, float2* __restrict__ partial, float wnorm, float wunnorm) {
    static_assert(TZ * TY * TX == 128 && 32 % (TY * TX) == 0 && TX % 2 == 0, "128-point tile, whole z rows per block");
    constexpr int VL = SPREAD_VL;
    constexpr int EB = SPIPE_EB;
    constexpr int ZR = 32 / (TY * TX);
    static_assert(EB == 16, "one MFMA K block per entry batch");
    __shared__ __attribute__((aligned(16))) float4 s_rec0[EB][SPIPE_RP];
    __shared__ __attribute__((aligned(16))) float4 s_rec1[EB][SPIPE_RP];
    __shared__ __attribute__((aligned(16))) float2 s_val0[EB][VL];
    __shared__ __attribute__((aligned(16))) float2 s_val1[EB][VL];
    __shared__ __attribute__((aligned(16))) float2 s_val2[EB][VL];
    __shared__ __attribute__((aligned(16))) float s_w[EB][TZ + TY + TX];  // 16-byte aligned rows
    __shared__ int s_id0[EB];
    __shared__ int s_id1[EB];

    const int nvch = B.C / VL;
    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int4 it = items[bid / nvch];
    const int vch = bid % nvch;
    const int lane = threadIdx.x;
    const int h = lane >> 5;     // K half: entries 8h .. 8h + 7
    const int col = lane & 31;   // A row / B and D column
    // this lane's A row: tile point (zl, yl, xl) within a row block
    const int zl = col / (TY * TX), yl = (col / TX) % TY, xl = col % TX;

    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tix = (int)(tile % g.ntiles[2]);
    const int tiy = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int oz = tiz * g.tile[0], oy = tiy * g.tile[1], ox = tix * g.tile[2];
    const int ez = min(g.tile[0], g.os[0] - oz), ey = min(g.tile[1], g.os[1] - oy);
    const int ex = min(g.tile[2], g.os[2] - ox);
    const int rec4 = g.rec / 4;
    const char* ybase = reinterpret_cast<const char*>(ys);

    pf_f16 acc[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[mb][nb][r] = 0.f;
    int E = 0;  // the tile's values are scaled by 2^-E

    // one batch (values in val, weights in s_w) into the accumulators
    auto accumulate = [&](const float2 (&val)[EB][VL]) {
        // B operand: entry 8h + j, column col of value block nb (value
        // nb * 16 + col / 2, real or imaginary part)
        float bv[2][8];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int j = 0; j < 8; ++j) bv[nb][j] = reinterpret_cast<const float*>(&val[8 * h + j][nb * 16])[col];
        pf_half8 bhi[2], blo[2];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            pf_half2 hi[4], lo[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                smfma_split(ldexpf(bv[nb][2 * q], -E), ldexpf(bv[nb][2 * q + 1], -E), hi[q], lo[q]);
            bhi[nb] = smfma_cat(hi[0], hi[1], hi[2], hi[3]);
            blo[nb] = smfma_cat(lo[0], lo[1], lo[2], lo[3]);
        }
        // A operand, one row block at a time (its operands live only for its
        // 6 products): entry 8h + j at this lane's point of row block mb
        float wyx[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) wyx[j] = s_w[8 * h + j][TZ + yl] * s_w[8 * h + j][TZ + TY + xl];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            pf_half2 hi[4], lo[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                smfma_split(s_w[8 * h + 2 * q][mb * ZR + zl] * wyx[2 * q],
                            s_w[8 * h + 2 * q + 1][mb * ZR + zl] * wyx[2 * q + 1], hi[q], lo[q]);
            const pf_half8 ahi = smfma_cat(hi[0], hi[1], hi[2], hi[3]);
            const pf_half8 alo = smfma_cat(lo[0], lo[1], lo[2], lo[3]);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bhi[nb], acc[mb][nb], 0, 0, 0);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, blo[nb], acc[mb][nb], 0, 0, 0);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bhi[nb], acc[mb][nb], 0, 0, 0);
        }
    };

    // batch b covers entries [b * EB, min(cnt, b * EB + EB)); its records go
    // to record slot b % 2, its values to value slot b % 3, and the copies of
    // batch b bring the sample ids of batch b + 2 into id slot b % 2.  The
    // loop's only vector memory instructions are these copies, SMFMA_NCOPY
    // per batch: when batch b is due, exactly the copies of batch b + 1 (if
    // it exists) were issued after its own.
    const int cnt = it.z - it.y;
    if (cnt > 0) {
        const int nbt = (cnt + EB - 1) / EB;
        const int32_t* sl = slist + it.y;
        // batch b's 16 ids (batches past the last reread the last one;
        // entries past the item's end come from the next item or the list's
        // padding and are never used)
        auto ids_src = [&](int b) { return sl + (int64_t)min(b, nbt - 1) * EB; };
        // every slot is named statically at each use, so the compiler can
        // tell the copies into one slot from reads of another and waits only
        // where this code does
        auto issue = [&](int b, int sid, float4 (&rec)[EB][SPIPE_RP], float2 (&val)[EB][VL], int (&ids)[EB]) {
            smfma_issue(rec, val, ids, ids_src(b + 2), sid, min(EB, cnt - b * EB), rec4, taprec, ybase, B.C, vch);
        };
        // batch b: records in rec (slot b % 2, which batch b + 2's records
        // then reuse), values in cur (slot b % 3), ids of batch b + 2 in ids
        // (slot b % 2, reused for those of batch b + 4); batch b + 2's values
        // go to nxt (slot (b + 2) % 3).  One wave per block: no block barrier
        // (it would also wait for the copies in flight), the compiler fences
        // keep LDS reads after the waits.
        auto step = [&](int b, float4 (&rec)[EB][SPIPE_RP], float2 (&nxt)[EB][VL], const float2 (&cur)[EB][VL],
                        int (&ids)[EB]) {
            if (b + 1 < nbt)
                __builtin_amdgcn_s_waitcnt(0xF70 | SMFMA_NCOPY);  // vmcnt(NCOPY); lgkm/exp not waited on
            else
                __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
            asm volatile("" ::: "memory");
            smfma_weights<TZ, TY, TX>(rec, min(EB, cnt - b * EB), g, oz, oy, ox, wnorm, s_w);
            if (b + 2 < nbt) {
                const int sid = ids[lane & (EB - 1)];
                asm volatile("" ::: "memory");
                issue(b + 2, sid, rec, nxt, ids);
            }
            asm volatile("" ::: "memory");
            accumulate(cur);
        };
        const int sid0 = sl[lane & (EB - 1)];
        const int sid1 = sl[min(EB + (lane & (EB - 1)), cnt - 1)];
        issue(0, sid0, s_rec0, s_val0, s_id0);
        if (nbt > 1) issue(1, sid1, s_rec1, s_val1, s_id1);
        // the largest |value| of this item's samples in this chunk
        frexpf(tmax[bid], &E);  // tmax < 2^E
        E -= 13;                // scaled values below 2^13 (fp16 holds up to 65504)
        // six steps per pass (slots repeat every 6 batches), one loop exit
        for (int b = 0; b < nbt; b += 6) {
            step(b, s_rec0, s_val2, s_val0, s_id0);
            if (b + 1 < nbt) step(b + 1, s_rec1, s_val0, s_val1, s_id1);
            if (b + 2 < nbt) step(b + 2, s_rec0, s_val1, s_val2, s_id0);
            if (b + 3 < nbt) step(b + 3, s_rec1, s_val2, s_val0, s_id1);
            if (b + 4 < nbt) step(b + 4, s_rec0, s_val0, s_val1, s_id0);
            if (b + 5 < nbt) step(b + 5, s_rec1, s_val1, s_val2, s_id1);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // no copy outlives the wave's LDS

    // accumulator element r of block (mb, nb) is tile row i = 8 (r / 4) +
    // 4 h + r % 4 of block mb (z = mb * ZR + i / (TY * TX), y = (i / TX) % TY,
    // x = i % TX), real column col of block nb.  Rows q = r / 4 hold 4
    // consecutive x of one (z, y) tile row.
    const int E0 = E;
    const int part = col & 1;  // 0: real, 1: imaginary
    const bool to_grid = it.w < 0;
    const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
    const int zmax = to_grid ? ez : Tz, ymax = to_grid ? ey : Ty, xmax = to_grid ? ex : Tx;
    float* base[2];        // the lane's value at tile point (0, 0, 0)
    int64_t sz, sy, sx;    // float strides of the tile's z, y, x
    if (to_grid) {
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const int u = vch * VL + nb * 16 + (col >> 1);
            const int64_t slab = (int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + u / B.Cv;
            const int64_t p0 = ((int64_t)oz * g.os[1] + oy) * g.os[2] + ox;
            base[nb] = reinterpret_cast<float*>(grid + (slab * g.gtot + p0) * B.Cv + u % B.Cv) + part;
        }
        sx = 2 * (int64_t)B.Cv;
        sy = sx * g.os[2];
        sz = sy * g.os[1];
    } else {
        float* d = reinterpret_cast<float*>(partial + ((int64_t)it.w * nvch + vch) * (int64_t)(Tx * Ty * Tz) * VL +
                                            (col >> 1)) + part;
        base[0] = d;
        base[1] = d + 32;  // values 16 .. 31 of the chunk
        sx = 2 * VL;
        sy = sx * Tx;
        sz = sy * Ty;
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i0 = 8 * q + 4 * h;  // first row of the 4 (x = i0 % TX .. + 3)
            const int lz = mb * ZR + i0 / (TY * TX), ly = (i0 / TX) % TY, lx0 = i0 % TX;
            if (lz >= zmax || ly >= ymax) continue;
            const int64_t off = lz * sz + ly * sy + lx0 * sx;
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                float* row = base[nb] + off;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (lx0 + e < xmax) row[e * sx] = ldexpf(acc[mb][nb][4 * q + e], E0) * wunnorm;
            }
        }
}

// -------------------------------------------------------------------------
// The same software-pipelined spread for whole 8- or 16-value chunks (plain
// batches of 8 or 16 coils, or multiples): VL values per wave, parts split
// x (and z where the tile is narrow) as in k_spread_reg, the values of an
// entry copied by VL / 2 lanes x 16 bytes.  The 32-value kernel above is kept
// as written: this generalised form compiles it to a schedule measured 4 %
// slower on the 3-D headline.
// -------------------------------------------------------------------------
template <int TZ, int TY, int TX>
constexpr int spipe_row_s() { return (TZ + TY + TX + 2 + 3) / 4 * 4; }

// copies of one batch: each entry's tap record (rec4 pieces of 16 bytes; the
// slot's unused pieces take a harmless copy of piece 0) and its VL values;
// sid holds lane j's sample id for entries j < n
template <int VL>
__device__ __forceinline__ void spipe_issue_s(float4 (&s_rec)[SPIPE_EB][SPIPE_RP], float2 (&val)[SPIPE_EB][VL],
                                            int sid, int n, int rec4, const float* __restrict__ taprec,
                                            const char* __restrict__ ybase, int C, int vch) {
    const int lane = threadIdx.x;
    const int c = lane & (SPIPE_RP - 1);
    const int cs = c < rec4 ? c : 0;
#pragma unroll
    for (int k = 0; k < SPIPE_EB * SPIPE_RP / 64; ++k) {
        const int e = k * (64 / SPIPE_RP) + (lane >> 3);
        const int se = __shfl(sid, e < n ? e : 0);
        const float* src = taprec + (int64_t)se * (rec4 * 4) + 4 * cs;
        if (k * (64 / SPIPE_RP) < n)
            __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)&s_rec[k * (64 / SPIPE_RP)][0],
                                             16, 0, 0);
    }
    // values: VL / 2 lanes x 16 bytes per entry, 128 / VL entries per copy
    constexpr int LPE = VL / 2, EPI = 64 / LPE;
    const int sub = lane / LPE, p16 = lane & (LPE - 1);
#pragma unroll
    for (int q = 0; q < SPIPE_EB / EPI; ++q) {
        const int j = q * EPI + sub;
        const unsigned sj = (unsigned)__shfl(sid, j < n ? j : 0);
        const unsigned off = (sj * (unsigned)C + (unsigned)(vch * VL)) * 8u + (unsigned)p16 * 16u;
        if (q * EPI < n)
            __builtin_amdgcn_global_load_lds(ybase + off, (__attribute__((address_space(3))) void*)&val[q * EPI][0],
                                             16, 0, 0);
    }
}

// one entry per lane: its record -> its weights at this tile's positions
// (0 outside the footprint) and the tile rows its footprint covers
template <int TZ, int TY, int TX>
__device__ __forceinline__ void spipe_weights_s(const float4 (&s_rec)[SPIPE_EB][SPIPE_RP], int n, const GeoParams& g,
                                              int oz, int oy, int ox,
                                              float (&s_w)[SPIPE_EB][spipe_row_s<TZ, TY, TX>()]) {
    const int lane = threadIdx.x;
    if (lane >= n) return;
    const float* rec_l = reinterpret_cast<const float*>(&s_rec[lane][0]);
    const int* ri = reinterpret_cast<const int*>(rec_l);
    int bz = oz - ri[0];
    bz += bz < 0 ? g.os[0] : 0;
    int by = oy - ri[1];
    by += by < 0 ? g.os[1] : 0;
    int bx = ox - ri[2];
    bx += bx < 0 ? g.os[2] : 0;
    unsigned mz = 0, my = 0;
#pragma unroll
    for (int a = 0; a < TZ; ++a) {
        int i = bz + a;
        i -= i >= g.os[0] ? g.os[0] : 0;
        const bool in = i < g.width[0];
        s_w[lane][a] = in ? rec_l[g.woff[0] + i] : 0.f;
        mz |= in ? 1u << a : 0u;
    }
#pragma unroll
    for (int b = 0; b < TY; ++b) {
        int i = by + b;
        i -= i >= g.os[1] ? g.os[1] : 0;
        const bool in = i < g.width[1];
        s_w[lane][TZ + b] = in ? rec_l[g.woff[1] + i] : 0.f;
        my |= in ? 1u << b : 0u;
    }
#pragma unroll
    for (int x = 0; x < TX; ++x) {
        int i = bx + x;
        i -= i >= g.os[2] ? g.os[2] : 0;
        s_w[lane][TZ + TY + x] = i < g.width[2] ? rec_l[g.woff[2] + i] : 0.f;
    }
    constexpr int NW = TZ + TY + TX;
    s_w[lane][NW] = __uint_as_float(mz);
    s_w[lane][NW + 1] = __uint_as_float(my);
}

// the batch's n entries into the lane's accumulators, in entry order
template <int TZ, int TY, int TXL, int VL, int ZP>
__device__ __forceinline__ void spipe_accumulate_s(const float2 (&val_b)[SPIPE_EB][VL], int n,
                                                 const float (&s_w)[SPIPE_EB][spipe_row_s<TZ, TY, 64 / (VL * ZP) * TXL>()],
                                                 float2 (&acc)[TZ / ZP][TY][TXL]) {
    constexpr int XP = 64 / (VL * ZP), TZL = TZ / ZP;
    const int lane = threadIdx.x;
    const int v = lane & (VL - 1);
    const int part = lane / VL;
    const int half = part & (XP - 1);         // x part
    const int zp = ZP == 1 ? 0 : part / XP;   // z part
    constexpr bool WR = spread_wreg<100 * TZ + TY>();
    constexpr bool VX = TXL <= 4;
    for (int j = 0; j < n; ++j) {
        constexpr int NW = TZ + TY + XP * TXL;
        const unsigned mz = __builtin_amdgcn_readfirstlane(__float_as_uint(s_w[j][NW]));
        const unsigned my = __builtin_amdgcn_readfirstlane(__float_as_uint(s_w[j][NW + 1]));
        const float2 val = val_b[j][v];
        const float* w = s_w[j];
        float2 vxk[TXL];
        float wxl[TXL];
#pragma unroll
        for (int k = 0; k < TXL; ++k) {
            const float wx = w[TZ + TY + half * TXL + k];
            if constexpr (VX) vxk[k] = make_float2(wx * val.x, wx * val.y);
            else wxl[k] = wx;
        }
        float wzr[WR ? TZ : 1], wyr[WR ? TY : 1];
        if constexpr (WR) {
#pragma unroll
            for (int a = 0; a < TZ; ++a) wzr[a] = w[a];
#pragma unroll
            for (int b = 0; b < TY; ++b) wyr[b] = w[TZ + b];
        }
#pragma unroll
        for (int a = 0; a < TZL; ++a) {
            float wz;
            if constexpr (ZP == 1) {
                if (!((mz >> a) & 1u)) continue;
                wz = WR ? wzr[WR ? a : 0] : w[a];
            } else {
                // the row pass runs when any z part's row a is inside the
                // footprint; a lane whose own row is outside reads weight 0
                unsigned rows = 0;
#pragma unroll
                for (int zq = 0; zq < ZP; ++zq) rows |= mz >> (zq * TZL + a);
                if (!(rows & 1u)) continue;
                if constexpr (WR) {
                    wz = wzr[a];
#pragma unroll
                    for (int zq = 1; zq < ZP; ++zq) wz = zp == zq ? wzr[WR ? zq * TZL + a : 0] : wz;
                } else {
                    wz = w[zp * TZL + a];
                }
            }
#pragma unroll
            for (int b = 0; b < TY; ++b) {
                if (!((my >> b) & 1u)) continue;
                const float wzy = wz * (WR ? wyr[WR ? b : 0] : w[TZ + b]);
                if constexpr (VX) {
#pragma unroll
                    for (int k = 0; k < TXL; ++k) {
                        acc[a][b][k].x += wzy * vxk[k].x;
                        acc[a][b][k].y += wzy * vxk[k].y;
                    }
                } else {
                    const float vx = wzy * val.x, vy = wzy * val.y;
#pragma unroll
                    for (int k = 0; k < TXL; ++k) {
                        acc[a][b][k].x += wxl[k] * vx;
                        acc[a][b][k].y += wxl[k] * vy;
                    }
                }
            }
        }
    }
}

template <int TZ, int TY, int TXL, int VL, int ZP>
__global__ void __launch_bounds__(64, spread_wreg<100 * TZ + TY>() ? 3 : 1)
    k_spread_pipe_s(GeoParams g, BatchParams B, const int4* __restrict__ items, const int32_t* __restrict__ slist,
                  const float* __restrict__ taprec, const float2* __restrict__ ys, float2* __restrict__ grid,
                  float2* __restrict__ partial) {
    constexpr int EB = SPIPE_EB;
    constexpr int XP = 64 / (VL * ZP), TZL = TZ / ZP, TX = XP * TXL;
    __shared__ __attribute__((aligned(16))) float4 s_rec[EB][SPIPE_RP];
    __shared__ __attribute__((aligned(16))) float2 s_val0[EB][VL];
    __shared__ __attribute__((aligned(16))) float2 s_val1[EB][VL];
    __shared__ __attribute__((aligned(16))) float s_w[EB][spipe_row_s<TZ, TY, TX>()];

    const int nvch = B.C / VL;  // plain values: one per coil
    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int4 it = items[bid / nvch];
    const int vch = bid % nvch;
    const int lane = threadIdx.x;
    const int v = lane & (VL - 1);
    const int part = lane / VL;
    const int half = part & (XP - 1);         // x part
    const int zp = ZP == 1 ? 0 : part / XP;   // z part
    const int u = vch * VL + v;

    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tix = (int)(tile % g.ntiles[2]);
    const int tiy = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tiz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int oz = tiz * g.tile[0], oy = tiy * g.tile[1], ox = tix * g.tile[2];
    const int ez = min(g.tile[0], g.os[0] - oz), ey = min(g.tile[1], g.os[1] - oy);
    const int ex = min(g.tile[2], g.os[2] - ox);
    const int rec4 = g.rec / 4;
    const char* ybase = reinterpret_cast<const char*>(ys);

    float2 acc[TZL][TY][TXL];
#pragma unroll
    for (int a = 0; a < TZL; ++a)
#pragma unroll
        for (int b = 0; b < TY; ++b)
#pragma unroll
            for (int k = 0; k < TXL; ++k) acc[a][b][k] = make_float2(0.f, 0.f);

    // batch i covers entries [it.y + i*EB, ...): its records land in s_rec,
    // its values in s_val0 (even i) or s_val1 (odd i).  Per batch: wait for
    // its copies, turn records into weights, start the next batch's copies
    // (records overwrite s_rec, which is no longer read), accumulate.
    const int cnt = it.z - it.y;
    const int32_t* sl = slist + it.y;  // the item's entries
    int sidA = lane < cnt ? sl[lane] : 0;
    int sidB = lane < cnt - EB ? sl[EB + lane] : 0;
    spipe_issue_s<VL>(s_rec, s_val0, sidA, min(EB, cnt), rec4, taprec, ybase, B.C, vch);
    for (int e0 = 0; e0 < cnt; e0 += 2 * EB) {
        {
            const int n = min(EB, cnt - e0);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            spipe_weights_s<TZ, TY, TX>(s_rec, n, g, oz, oy, ox, s_w);
            __syncthreads();
            const int n1 = min(EB, cnt - e0 - EB);
            if (n1 > 0) {
                // ids first: a reload of a spilled register before them waits on
                // nothing but itself
                const int sidC = lane < cnt - e0 - 2 * EB ? sl[e0 + 2 * EB + lane] : 0;
                spipe_issue_s<VL>(s_rec, s_val1, sidB, n1, rec4, taprec, ybase, B.C, vch);
                sidA = sidC;
            }
            spipe_accumulate_s<TZ, TY, TXL, VL, ZP>(s_val0, n, s_w, acc);
        }
        if (e0 + EB >= cnt) break;
        {
            const int n = min(EB, cnt - e0 - EB);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            spipe_weights_s<TZ, TY, TX>(s_rec, n, g, oz, oy, ox, s_w);
            __syncthreads();
            const int n2 = min(EB, cnt - e0 - 2 * EB);
            if (n2 > 0) {
                const int sidC = lane < cnt - e0 - 3 * EB ? sl[e0 + 3 * EB + lane] : 0;
                spipe_issue_s<VL>(s_rec, s_val0, sidA, n2, rec4, taprec, ybase, B.C, vch);
                sidB = sidC;
            }
            spipe_accumulate_s<TZ, TY, TXL, VL, ZP>(s_val1, n, s_w, acc);
        }
    }

    if (it.w < 0) {
        const int64_t slab = (int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + u / B.Cv;
        float2* base = grid + slab * g.gtot * B.Cv + u % B.Cv;
#pragma unroll
        for (int a = 0; a < TZL; ++a)
#pragma unroll
            for (int b = 0; b < TY; ++b)
#pragma unroll
                for (int k = 0; k < TXL; ++k) {
                    const int lx = half * TXL + k;
                    const int lz = zp * TZL + a;
                    if (lz < ez && b < ey && lx < ex) {
                        const int64_t gpt = ((int64_t)(oz + lz) * g.os[1] + (oy + b)) * g.os[2] + (ox + lx);
                        base[gpt * B.Cv] = acc[a][b][k];
                    }
                }
    } else {
        const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
        float2* dst = partial + ((int64_t)it.w * nvch + vch) * (int64_t)(Tx * Ty * Tz) * VL + v;
#pragma unroll
        for (int a = 0; a < TZL; ++a)
#pragma unroll
            for (int b = 0; b < TY; ++b)
#pragma unroll
                for (int k = 0; k < TXL; ++k) {
                    const int lx = half * TXL + k;
                    const int lz = zp * TZL + a;
                    if (lz < Tz && b < Ty && lx < Tx) dst[(int64_t)((lz * Ty + b) * Tx + lx) * VL] = acc[a][b][k];
                }
    }
}

// Sample permutation between the user layout (C, F*npts) and the tile-sorted
// layout (Mtot, C), through an LDS transpose of PSB consecutive USER samples
// x up to PSC batch values per block: the user side reads/writes contiguous
// runs of each batch plane, the sorted side whole coil rows (C values of one
// sample are contiguous there) at iperm[j].
constexpr int PSB = 64;
constexpr int PSC = 64;

// user -> sorted, times the density weight
__global__ void __launch_bounds__(TB) k_sort(int64_t mtot, int C, const int32_t* __restrict__ iperm,
                                             const float2* __restrict__ y, const float* __restrict__ w,
                                             float2* __restrict__ ys) {
    __shared__ float2 t[PSB][PSC + 1];
    const int64_t j0 = (int64_t)blockIdx.x * PSB;
    const int c0 = blockIdx.y * PSC;
    const int cc = min(PSC, C - c0);
    const int nj = (int)min<int64_t>(PSB, mtot - j0);
    for (int i = threadIdx.x; i < cc * PSB; i += TB) {
        const int c = i / PSB, jl = i % PSB;
        if (jl < nj) {
            float2 v = y[(int64_t)(c0 + c) * mtot + j0 + jl];
            if (w) {
                const float a = w[j0 + jl];
                v.x *= a;
                v.y *= a;
            }
            t[jl][c] = v;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nj * cc; i += TB) {
        const int jl = i / cc, c = i % cc;
        ys[(int64_t)iperm[j0 + jl] * C + c0 + c] = t[jl][c];
    }
}

// sorted -> user, times factor and the density weight
__global__ void __launch_bounds__(TB) k_unsort(int64_t mtot, int C, const int32_t* __restrict__ iperm,
                                               const float2* __restrict__ ys, const float* __restrict__ w,
                                               float factor, float2* __restrict__ y) {
    __shared__ float2 t[PSB][PSC + 1];
    const int64_t j0 = (int64_t)blockIdx.x * PSB;
    const int c0 = blockIdx.y * PSC;
    const int cc = min(PSC, C - c0);
    const int nj = (int)min<int64_t>(PSB, mtot - j0);
    for (int i = threadIdx.x; i < nj * cc; i += TB) {
        const int jl = i / cc, c = i % cc;
        t[jl][c] = ys[(int64_t)iperm[j0 + jl] * C + c0 + c];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cc * PSB; i += TB) {
        const int c = i / PSB, jl = i % PSB;
        if (jl < nj) {
            const float a = factor * (w ? w[j0 + jl] : 1.0f);
            const float2 v = t[jl][c];
            y[(int64_t)(c0 + c) * mtot + j0 + jl] = make_float2(v.x * a, v.y * a);
        }
    }
}

// 16-byte forms of the two permutations (even C, even Mtot, 16-byte aligned
// arrays): a block moves QSB samples x QSC values, every thread issues all of
// its QN loads before it stores anything, the sorted side as 16-byte pieces of
// whole coil rows, the user side as pairs of consecutive samples.
constexpr int QSB = 128;
constexpr int QSC = 32;
constexpr int QN = QSB * QSC / 2 / TB;  // float4 per thread and phase

// user -> sorted, times the density weight
// With smax (C a multiple of 32, one 32-value chunk per block row): also
// the chunk's max |re|, |im| per sorted sample, smax[t * C / 32 + chunk], for
// the matrix-core spread's value scale.
__global__ void __launch_bounds__(TB) k_sort4(int64_t mtot, int C, const int32_t* __restrict__ iperm,
                                              const float2* __restrict__ y, const float* __restrict__ w,
                                              float2* __restrict__ ys, float* __restrict__ smax) {
    __shared__ float2 t[QSB][QSC + 1];
    const int64_t j0 = (int64_t)blockIdx.x * QSB;
    const int c0 = blockIdx.y * QSC;
    const int cc = min(QSC, C - c0);
    const int nj = (int)min<int64_t>(QSB, mtot - j0);  // even
    float4 v[QN];
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int c = i / (QSB / 2), jl = 2 * (i % (QSB / 2));
        if (c < cc && jl < nj) v[k] = *reinterpret_cast<const float4*>(y + (int64_t)(c0 + c) * mtot + j0 + jl);
    }
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int c = i / (QSB / 2), jl = 2 * (i % (QSB / 2));
        if (c < cc && jl < nj) {
            const float a0 = w ? w[j0 + jl] : 1.0f, a1 = w ? w[j0 + jl + 1] : 1.0f;
            t[jl][c] = make_float2(v[k].x * a0, v[k].y * a0);
            t[jl + 1][c] = make_float2(v[k].z * a1, v[k].w * a1);
        }
    }
    __syncthreads();
    const int h = cc >> 1;  // float4 per sorted row
    int row[QN];
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int jl = i / h;
        row[k] = jl < nj ? iperm[j0 + jl] : 0;
    }
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int jl = i / h, q = i - jl * h;
        float m = 0.f;
        if (jl < nj) {
            const float2 a = t[jl][2 * q], b = t[jl][2 * q + 1];
            *reinterpret_cast<float4*>(ys + (int64_t)row[k] * C + c0 + 2 * q) = make_float4(a.x, a.y, b.x, b.y);
            m = fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(b.x), fabsf(b.y)));
        }
        if (smax) {  // h = 16: a row's pieces are one 16-lane DPP row
            int x = __float_as_int(m);
            x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));  // row_shr:1
            x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));  // row_shr:2
            x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));  // row_shr:4
            x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));  // row_shr:8
            if (q == 15 && jl < nj) smax[(int64_t)row[k] * (C / QSC) + blockIdx.y] = __int_as_float(x);
        }
    }
}

// user -> sorted with the reduction entries expanded (the fused adjoint with B0
// segments or a subspace basis, all batch items in one vector): sorted row
// iperm[j] gets the V = R*C values y[c, j] * w[j] * conj(coef(r)) in the
// grid's value order r*C + c, coef(r = l*K + k) = Bt[l, t(j)] * basis[f(j), k]
// as the on-the-fly spread forms them (sample_coef), so the spread sees a
// plain batch of V items.  A block takes 2^qsh consecutive samples: coil-major
// loads on consecutive samples, a transpose through LDS, the samples'
// coefficients and destination rows in LDS, then whole sorted rows written
// with 16-byte stores (pairs of values of one r; C even).  Divisions by V/2,
// C and R are (n * ceil(2^32 / d)) >> 32, exact for n * d < 2^32 (the
// launcher bounds V by 1024 and a block's samples by 32).
__device__ __forceinline__ int udiv_m(int n, uint64_t m) { return (int)(((uint64_t)(unsigned)n * m) >> 32); }

template <bool PAIR>
__global__ void __launch_bounds__(TB) k_sort_x(GeoParams g, BatchParams B, int64_t mtot, int qsh, uint64_t m_out,
                                               uint64_t m_c, uint64_t m_r, const int32_t* __restrict__ iperm,
                                               const float* __restrict__ taprec, const float2* __restrict__ y,
                                               const float* __restrict__ w, float2* __restrict__ ys) {
    extern __shared__ float2 sm[];
    const int QX = 1 << qsh;
    const int C = B.C, CP = B.C + 1, R = B.R, V = B.R * B.C;
    float2* t = sm;                                       // [QX][C + 1] weighted values
    float2* cf = sm + QX * CP;                            // [QX][R] coefficients
    int* srow = reinterpret_cast<int*>(cf + QX * R);      // [QX] sorted rows
    const int64_t j0 = (int64_t)blockIdx.x << qsh;
    const int nj = (int)min<int64_t>(QX, mtot - j0);
    for (int i = threadIdx.x; i < QX * C; i += TB) {
        const int c = i >> qsh, jl = i & (QX - 1);
        if (jl < nj) {
            const float a = w ? w[j0 + jl] : 1.0f;
            const float2 v = y[(int64_t)c * mtot + j0 + jl];
            t[jl * CP + c] = make_float2(v.x * a, v.y * a);
        }
    }
    for (int i = threadIdx.x; i < nj * R; i += TB) {
        const int jl = udiv_m(i, m_r), r = i - jl * R;
        const int64_t j = j0 + jl;
        const int s = iperm[j];
        if (r == 0) srow[jl] = s;
        const int f = reinterpret_cast<const int*>(taprec + (int64_t)s * g.rec)[3];
        const int64_t tread = B.b0t ? (j % g.npts) % B.nt : 0;
        cf[jl * R + r] = sample_coef(B, r, f, tread);
    }
    __syncthreads();
    if (PAIR) {
        const int VH = V >> 1;
        for (int i = threadIdx.x; i < nj * VH; i += TB) {
            const int jl = udiv_m(i, m_out), e = 2 * (i - jl * VH);
            const int r = udiv_m(e, m_c), c = e - r * C;
            const float2 k = cf[jl * R + r];
            const float2 a = cmulc(t[jl * CP + c], k), b = cmulc(t[jl * CP + c + 1], k);
            *reinterpret_cast<float4*>(ys + (int64_t)srow[jl] * V + e) = make_float4(a.x, a.y, b.x, b.y);
        }
    } else {
        for (int i = threadIdx.x; i < nj * V; i += TB) {
            const int jl = udiv_m(i, m_out), e = i - jl * V;
            const int r = udiv_m(e, m_c), c = e - r * C;
            ys[(int64_t)srow[jl] * V + e] = cmulc(t[jl * CP + c], cf[jl * R + r]);
        }
    }
}

// sorted -> user, times factor and the density weight
__global__ void __launch_bounds__(TB) k_unsort4(int64_t mtot, int C, const int32_t* __restrict__ iperm,
                                                const float2* __restrict__ ys, const float* __restrict__ w,
                                                float factor, float2* __restrict__ y) {
    __shared__ float2 t[QSB][QSC + 1];
    const int64_t j0 = (int64_t)blockIdx.x * QSB;
    const int c0 = blockIdx.y * QSC;
    const int cc = min(QSC, C - c0);
    const int h = cc >> 1;
    const int nj = (int)min<int64_t>(QSB, mtot - j0);  // even
    int row[QN];
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int jl = i / h;
        row[k] = jl < nj ? iperm[j0 + jl] : 0;
    }
    float4 v[QN];
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int jl = i / h, q = i - jl * h;
        if (jl < nj) v[k] = *reinterpret_cast<const float4*>(ys + (int64_t)row[k] * C + c0 + 2 * q);
    }
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int jl = i / h, q = i - jl * h;
        if (jl < nj) {
            t[jl][2 * q] = make_float2(v[k].x, v[k].y);
            t[jl][2 * q + 1] = make_float2(v[k].z, v[k].w);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < QN; ++k) {
        const int i = threadIdx.x + k * TB;
        const int c = i / (QSB / 2), jl = 2 * (i % (QSB / 2));
        if (c < cc && jl < nj) {
            const float a0 = factor * (w ? w[j0 + jl] : 1.0f), a1 = factor * (w ? w[j0 + jl + 1] : 1.0f);
            const float2 u0 = t[jl][c], u1 = t[jl + 1][c];
            *reinterpret_cast<float4*>(y + (int64_t)(c0 + c) * mtot + j0 + jl) =
                make_float4(u0.x * a0, u0.y * a0, u1.x * a1, u1.y * a1);
        }
    }
}

unsigned grid_stride_blocks(int64_t n) {
    int64_t b = (n + TB - 1) / TB;
    return (unsigned)std::min<int64_t>(b, 1 << 20);
}

}  // namespace

int pf_launch_deapod_pad(const pf_plan* p, const pf_batch& b, const float2* img, float2* grid, hipStream_t s) {
    GeoParams g = geo(p);
    BatchParams B = bparams(b);
    int64_t total = (int64_t)b.nfg * b.Cb * g.gtot * b.R * b.Cv;
    hipLaunchKernelGGL(k_deapod_pad, dim3(grid_stride_blocks(total)), dim3(TB), 0, s, g, B, p->apod, img,
                       grid);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_crop_deapod(const pf_plan* p, const pf_batch& b, const float2* grid, float2* img, float factor,
                          hipStream_t s) {
    GeoParams g = geo(p);
    BatchParams B = bparams(b);
    int64_t total = (int64_t)b.C * b.nimg * g.ntot;
    hipLaunchKernelGGL(k_crop_deapod, dim3(grid_stride_blocks(total)), dim3(TB), 0, s, g, B, p->apod,
                       grid, img, factor);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

using GatherLaunch = void (*)(dim3, hipStream_t, const GeoParams&, const BatchParams&, const pf_plan*, const float2*,
                              float2*);

template <int CPL, int WY, int WX, bool DEEP = false>
void launch_gather_w(dim3 grid, hipStream_t s, const GeoParams& g, const BatchParams& B, const pf_plan* p,
                     const float2* in, float2* ys) {
    hipLaunchKernelGGL((k_gather_w<CPL, WY, WX, DEEP>), grid, dim3(GTB), 0, s, g, B, p->mtot, p->taprec, p->perm,
                       in, ys);
}

template <int CPL, int W>
constexpr GatherLaunch gw_sq() { return &launch_gather_w<CPL, W, W>; }

template <int CPL, int W>
constexpr GatherLaunch gw_deep() { return &launch_gather_w<CPL, W, W, true>; }

// square y/x widths of 3-D plans: [cpl - 1][w - 1]
const GatherLaunch GATHER_W_DEEP[2][PF_WMAX] = {
    {nullptr, gw_deep<1, 2>(), gw_deep<1, 3>(), gw_deep<1, 4>(), gw_deep<1, 5>(), gw_deep<1, 6>(), gw_deep<1, 7>(),
     gw_deep<1, 8>()},
    {nullptr, gw_deep<2, 2>(), gw_deep<2, 3>(), gw_deep<2, 4>(), gw_deep<2, 5>(), gw_deep<2, 6>(), gw_deep<2, 7>(),
     gw_deep<2, 8>()},
};

// [cpl - 1][wy == 1 ? 0 : 1][wx - 1]
const GatherLaunch GATHER_W[2][2][PF_WMAX] = {
    {{&launch_gather_w<1, 1, 1>, &launch_gather_w<1, 1, 2>, &launch_gather_w<1, 1, 3>, &launch_gather_w<1, 1, 4>,
      &launch_gather_w<1, 1, 5>, &launch_gather_w<1, 1, 6>, &launch_gather_w<1, 1, 7>, &launch_gather_w<1, 1, 8>},
     {nullptr, gw_sq<1, 2>(), gw_sq<1, 3>(), gw_sq<1, 4>(), gw_sq<1, 5>(), gw_sq<1, 6>(), gw_sq<1, 7>(),
      gw_sq<1, 8>()}},
    {{&launch_gather_w<2, 1, 1>, &launch_gather_w<2, 1, 2>, &launch_gather_w<2, 1, 3>, &launch_gather_w<2, 1, 4>,
      &launch_gather_w<2, 1, 5>, &launch_gather_w<2, 1, 6>, &launch_gather_w<2, 1, 7>, &launch_gather_w<2, 1, 8>},
     {nullptr, gw_sq<2, 2>(), gw_sq<2, 3>(), gw_sq<2, 4>(), gw_sq<2, 5>(), gw_sq<2, 6>(), gw_sq<2, 7>(),
      gw_sq<2, 8>()}},
};

// Which gather instance a batch runs: the unrolled k_gather_w for square (or
// 1-D) y/x widths with 32-bit grid offsets (the DEEP two-wave instances for
// 3-D), else the generic k_gather; cpl = coils per lane.
// Batches the LDS-staged gather runs on: 3-D plans with gather lists (one
// width 3..8 on all axes), plain values (no basis, no field map), an even
// batch, and grid point indices within 31 bits.  Their forward grid is laid
// out coil-pair major (Cv = 2, Cb = C / 2).
bool pf_gather_col_layout(const pf_plan* p, const pf_batch& b) {
    const int w = p->width[2];
    return PF_GATHER_COL && p->gl.nitems > 0 && b.R == 1 && !b.basis && !b.b0t && b.C % 2 == 0 && w >= 3 && w <= 8 &&
           pf_gtot(p) < (1ll << 31);
}

using GColLaunch = void (*)(hipStream_t, const GeoParams&, const BatchParams&, const pf_plan*, const float2*, float2*);

template <int W>
void launch_gather_col(hipStream_t s, const GeoParams& g, const BatchParams& B, const pf_plan* p, const float2* in,
                       float2* ys) {
    hipLaunchKernelGGL((k_gather_col<W>), dim3((unsigned)p->gl.nitems), dim3(GCOL_TB), 0, s, g, B, p->gl.items,
                       p->gl.ids, p->taprec, in, ys, p->gl.ncol[0], p->gl.ncol[1], p->gl.nzw);
}

// [W - 3], W = 3 .. 8
const GColLaunch GATHER_COL[6] = {&launch_gather_col<3>, &launch_gather_col<4>, &launch_gather_col<5>,
                                  &launch_gather_col<6>, &launch_gather_col<7>, &launch_gather_col<8>};

struct GatherChoice {
    GatherLaunch fn = nullptr;  // null: generic k_gather<cpl>
    GColLaunch col = nullptr;   // LDS-staged gather (takes precedence)
    int cpl = 1;
    bool deep = false;
};

static GatherChoice gather_choice(const pf_plan* p, const pf_batch& b) {
    GatherChoice c;
    c.cpl = b.Cv % 2 == 0 ? 2 : 1;
    const int wy = p->width[1], wx = p->width[2];
    const bool idx32 = (int64_t)pf_gtot(p) * b.R * b.Cv < (1ll << 31);
    if (b.Cv == 2 && pf_gather_col_layout(p, b)) {
        c.col = GATHER_COL[wx - 3];
        c.cpl = 2;
        return c;
    }
    if (idx32 && (wy == 1 || wy == wx)) {
        c.deep = p->width[0] > 1 && wy == wx && wy > 1;
        c.fn = c.deep ? GATHER_W_DEEP[c.cpl - 1][wx - 1] : GATHER_W[c.cpl - 1][wy == 1 ? 0 : 1][wx - 1];
        if (!c.fn) c.deep = false;
    }
    return c;
}

std::string pf_gather_name(const pf_plan* p, const pf_batch& b) {
    const GatherChoice c = gather_choice(p, b);
    if (c.col) return "k_gather_col<" + std::to_string(p->width[2]) + ">";
    if (!c.fn) return "k_gather<" + std::to_string(c.cpl) + ">";
    return "k_gather_w<" + std::to_string(c.cpl) + "," + std::to_string(p->width[1]) + "," +
           std::to_string(p->width[2]) + "," + (c.deep ? "true" : "false") + ">";
}

int pf_launch_gather(const pf_plan* p, const pf_batch& b, const float2* grid, float2* ys, hipStream_t s) {
    GeoParams g = geo(p);
    BatchParams B = bparams(b);
    const GatherChoice c = gather_choice(p, b);
    if (c.col) {
        c.col(s, g, B, p, grid, ys);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (c.fn) {
        const int64_t threads = p->mtot * (b.C / c.cpl);
        c.fn(dim3(nblk(threads, GTB)), s, g, B, p, grid, ys);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (c.cpl == 2) {
        int64_t threads = p->mtot * (b.C / 2);
        hipLaunchKernelGGL(k_gather<2>, dim3(nblk(threads, TB)), dim3(TB), 0, s, g, B, p->mtot, p->taprec,
                           p->perm, grid, ys);
    } else {
        int64_t threads = p->mtot * b.C;
        hipLaunchKernelGGL(k_gather<1>, dim3(nblk(threads, TB)), dim3(TB), 0, s, g, B, p->mtot, p->taprec,
                           p->perm, grid, ys);
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

// Register-spread instances: tile (z, y, x) held by one wave of VL values
// x (64 / VL) x parts of TXL points; VL = 32 holds 64 complex accumulators
// per lane, the small-batch instances (VL = 16, 8) proportionally fewer.
using RegLaunch = void (*)(dim3, hipStream_t, const GeoParams&, const BatchParams&, const pf_slist&,
                           const pf_sitems&, const pf_plan*, const float2*, float2*, float2*, const float*);

// I0(x) by its power series (the normalisation of the matrix-core spread's
// weights needs the kernel's peak only roughly)
static double host_i0(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 200 && term > 1e-17 * sum; ++k) {
        const double f = 0.5 * x / k;
        term *= f * f;
        sum += term;
    }
    return sum;
}

// 1 / (peak of the separable kernel): the weights a tile's A operand holds
// are at most 1
static float spread_wnorm(const pf_plan* p) {
    double peak = 1.0;
    for (int a = 0; a < 3; ++a)
        if (p->width[a] > 1) peak *= host_i0((double)p->beta[a]);
    return (float)(1.0 / peak);
}

// the matrix-core spread's tiles: 128 points, whole z rows per 32-row block
template <int TZ, int TY, int TX>
constexpr bool smfma_tile() { return TZ * TY * TX == 128 && 32 % (TY * TX) == 0 && TX % 2 == 0; }

template <int TZ, int TY, int TXL, int VL, int ZP>
void launch_reg(dim3 grid, hipStream_t s, const GeoParams& g, const BatchParams& B, const pf_slist& L,
                const pf_sitems& I, const pf_plan* p, const float2* ys, float2* out, float2* partial,
                const float* tmax) {
    if (B.dma && PF_SPREAD_PIPE) {
        if constexpr (PF_SPREAD_MFMA && VL == SPREAD_VL && ZP == 1 && smfma_tile<TZ, TY, 2 * TXL>()) {
            const float wn = spread_wnorm(p);
            hipLaunchKernelGGL((k_spread_mfma<TZ, TY, 2 * TXL>), grid, dim3(64), 0, s, g, B, I.sitems, L.slist,
                               p->taprec, ys, tmax, out, partial, wn, 1.0f / wn);
        } else if constexpr (VL == SPREAD_VL && ZP == 1)
            hipLaunchKernelGGL((k_spread_pipe<TZ, TY, TXL>), grid, dim3(64), 0, s, g, B, I.sitems, L.slist,
                               p->taprec, ys, out, partial);
        else
            hipLaunchKernelGGL((k_spread_pipe_s<TZ, TY, TXL, VL, ZP>), grid, dim3(64), 0, s, g, B, I.sitems,
                               L.slist, p->taprec, ys, out, partial);
        return;
    }
    hipLaunchKernelGGL((k_spread_reg<TZ, TY, TXL, VL, ZP>), grid, dim3(64), 0, s, g, B, I.sitems, L.slist,
                       p->taprec, p->perm, ys, out, partial);
}

// the instance of tile (TZ, TY, TX) with VL values per wave: the 64 / VL
// parts split x, or x and z in two halves when the x extent is too narrow;
// null when neither fits
template <int TZ, int TY, int TX, int VL>
constexpr RegLaunch reg_fn() {
    if constexpr (TX * VL >= 64)
        return &launch_reg<TZ, TY, TX * VL / 64, VL, 1>;
    else if constexpr (TX * VL * 2 >= 64 && TZ % 2 == 0)
        return &launch_reg<TZ, TY, TX * VL * 2 / 64, VL, 2>;
    else
        return nullptr;
}

struct RegTile {
    int ndim, z, y, x;
    RegLaunch fn[3];  // VL = 32, 16, 8
};

template <int TZ, int TY, int TX>
constexpr RegTile reg_tile(int ndim) {
    return RegTile{ndim, TZ, TY, TX, {reg_fn<TZ, TY, TX, 32>(), reg_fn<TZ, TY, TX, 16>(), reg_fn<TZ, TY, TX, 8>()}};
}

const RegTile REG_TILES[] = {
    reg_tile<8, 4, 4>(3),  reg_tile<4, 4, 8>(3),  reg_tile<4, 8, 4>(3),  reg_tile<2, 4, 16>(3),
    reg_tile<1, 8, 16>(2), reg_tile<1, 16, 8>(2), reg_tile<1, 4, 32>(2), reg_tile<1, 32, 4>(2),
    reg_tile<1, 1, 64>(1),
};

// values per wave for V batch values: the smallest of 32, 16, 8 that holds
// them (32 beyond), among those the tile has an instance for
static int spread_vl_for(const RegTile& t, int V) {
    int want = V > 16 ? 0 : (V > 8 ? 1 : 2);
    while (want > 0 && !t.fn[want]) --want;
    return want;
}

struct RegChoice {
    const RegTile* tile = nullptr;
    int vl = 0;  // values per wave
    RegLaunch fn = nullptr;
};

// The smallest register instance that holds the plan's tile, or an empty
// choice (tiles larger than every instance use the LDS-staged spread).
static RegChoice spread_reg_choice(const pf_plan* p, const pf_batch& b) {
    RegChoice c;
    const int V = b.Cb * b.R * b.Cv;
    if (V < PF_SPREAD_REG_MIN) return c;
    for (const RegTile& t : REG_TILES) {
        if (t.ndim != p->ndim || p->tile[0] > t.z || p->tile[1] > t.y || p->tile[2] > t.x) continue;
        if (!c.tile || t.z * t.y * t.x < c.tile->z * c.tile->y * c.tile->x) c.tile = &t;
    }
    if (c.tile) {
        const int k = spread_vl_for(*c.tile, V);
        c.vl = 32 >> k;
        c.fn = c.tile->fn[k];
    }
    return c;
}

// Plain values (no basis, no B0) in whole chunks of the wave's VL coils: a
// wave's values of a sample are one contiguous run, copied global -> LDS with
// 32-bit offsets (sorted samples under 4 GiB) by the pipelined instances.
static bool spread_dma(const pf_plan* p, const pf_batch& b, const RegChoice& rc) {
    return !b.basis && !b.b0t && b.R == 1 && rc.vl > 0 && b.C % rc.vl == 0 &&
           (uint64_t)p->mtot * b.C * sizeof(float2) < (1ull << 32);
}

// the matrix-core spread runs for this batch (a 128-point register tile with
// whole 32-value chunks copied global -> LDS)
static bool spread_mfma(const pf_plan* p, const pf_batch& b) {
    const RegChoice rc = spread_reg_choice(p, b);
    if (!PF_SPREAD_MFMA || !PF_SPREAD_PIPE || !rc.fn || rc.vl != SPREAD_VL || !spread_dma(p, b, rc)) return false;
    const RegTile& t = *rc.tile;
    return t.x * rc.vl >= 64 && t.z * t.y * t.x == 128 && 32 % (t.y * t.x) == 0 && t.x % 2 == 0;
}

static size_t spread_tiles_bytes(const pf_plan* p, const pf_batch& b) {
    const int VC = pf_spread_coil_chunk(p, b);
    const int V = b.Cb * b.R * b.Cv;
    const int nvch = (V + VC - 1) / VC;
    const size_t tpts = (size_t)p->tile[0] * p->tile[1] * p->tile[2];
    return (size_t)pf_sitems_for(pf_slist_for(p, b), nvch).nslots * nvch * tpts * VC * sizeof(float2);
}

std::string pf_spread_name(const pf_plan* p, const pf_batch& b) {
    const RegChoice rc = spread_reg_choice(p, b);
    std::string name;
    if (!rc.fn) {
        name = "k_spread";
    } else {
        const RegTile& t = *rc.tile;
        const int zp = t.x * rc.vl >= 64 ? 1 : 2;
        const int txl = t.x * rc.vl * zp / 64;
        const std::string tile = std::to_string(t.z) + "," + std::to_string(t.y) + "," + std::to_string(txl);
        if (spread_mfma(p, b))
            name = "k_tile_max+k_spread_mfma<" + std::to_string(t.z) + "," + std::to_string(t.y) + "," +
                   std::to_string(t.x) + ">";
        else if (spread_dma(p, b, rc) && PF_SPREAD_PIPE)
            name = (rc.vl == SPREAD_VL && zp == 1) ? "k_spread_pipe<" + tile + ">"
                                                    : "k_spread_pipe_s<" + tile + "," + std::to_string(rc.vl) + "," +
                                                          std::to_string(zp) + ">";
        else
            name = "k_spread_reg<" + tile + "," + std::to_string(rc.vl) + "," + std::to_string(zp) + ">";
    }
    const int VC = pf_spread_coil_chunk(p, b);
    const int V = b.Cb * b.R * b.Cv;
    const int nvch = (V + VC - 1) / VC;
    if (pf_sitems_for(pf_slist_for(p, b), nvch).nritems > 0) name += "+k_spread_reduce";
    return name;
}

int pf_spread_coil_chunk(const pf_plan* p, const pf_batch& b) {
    const RegChoice rc = spread_reg_choice(p, b);
    if (rc.fn) return rc.vl;
    const int tpts = p->tile[0] * p->tile[1] * p->tile[2];
    const int V = b.Cb * b.R * b.Cv;
    // accumulator (2*VC+1 words per point) within the LDS budget, multiple of 4
    int vc = (SPREAD_ACC_BYTES / 4 / tpts - 1) / 2;
    vc = std::max(1, std::min(vc, 64));
    if (vc >= 4) vc &= ~3;
    return std::min(vc, V);
}

static size_t spread_lds_bytes(const pf_plan* p, int VC) {
    const size_t tpts = (size_t)p->tile[0] * p->tile[1] * p->tile[2];
    size_t words = tpts * (2 * VC + 1) + SB * 8 + SB * 3 * PF_WMAX;
    words += words % 2;
    return words * 4 + (size_t)SB * VC * sizeof(float2);
}

static size_t al256(size_t n) { return (n + 255) / 256 * 256; }

// [partial tiles | per-sample chunk maxima | per-item chunk maxima], the
// maxima for the matrix-core spread only
size_t pf_spread_partial_bytes(const pf_plan* p, const pf_batch& b) {
    size_t n = spread_tiles_bytes(p, b);
    if (spread_mfma(p, b)) {
        const int nvch = b.C / SPREAD_VL;
        n = al256(n) + al256((size_t)p->mtot * nvch * sizeof(float)) +
            (size_t)pf_sitems_for(pf_slist_for(p, b), nvch).nsitems * nvch * sizeof(float);
    }
    return n;
}

float* pf_spread_smax(const pf_plan* p, const pf_batch& b, float2* partial) {
    if (!partial || !spread_mfma(p, b)) return nullptr;
    return reinterpret_cast<float*>(reinterpret_cast<char*>(partial) + al256(spread_tiles_bytes(p, b)));
}

int pf_spread_reg_tiles(int ndim, int* tiles, int cap) {
    int n = 0;
    for (const RegTile& t : REG_TILES) {
        if (t.ndim != ndim) continue;
        if (n < cap && tiles) {
            tiles[3 * n] = t.z;
            tiles[3 * n + 1] = t.y;
            tiles[3 * n + 2] = t.x;
        }
        ++n;
    }
    return n;
}

int pf_launch_spread(const pf_plan* p, const pf_batch& b, const float2* ys, float2* grid, float2* partial,
                     hipStream_t s, bool smax_ready, bool live) {
    GeoParams g = geo(p);
    BatchParams B = bparams(b);
    const int VC = pf_spread_coil_chunk(p, b);
    const int V = b.Cb * b.R * b.Cv;
    const int nvch = (V + VC - 1) / VC;
    const size_t lds = spread_lds_bytes(p, VC);
    if (lds > 160 * 1024) return pf_fail(PF_ERR_UNSUPPORTED, "spread tile exceeds LDS");
    const pf_slist& L = pf_slist_for(p, b);
    pf_sitems I = pf_sitems_for(L, nvch);
    if (live && I.lsitems) {  // the same items without the empty tiles of uncovered columns
        I.sitems = I.lsitems;
        I.nsitems = I.nlsitems;
    }
    if (I.nslots > 0 && !partial) return pf_fail(PF_ERR_ARG, "spread needs a partial-tile workspace");
    const int64_t nb = (int64_t)I.nsitems * nvch;
    if (nb > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many spread blocks");
    const int nsub = (int)(((int64_t)p->tile[0] * p->tile[1] * p->tile[2] * VC + TB - 1) / TB);
    if ((int64_t)I.nritems * nvch * nsub > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many reduce blocks");
    const RegChoice rc = spread_reg_choice(p, b);
    if (rc.fn) {
        B.dma = spread_dma(p, b, rc) ? 1 : 0;
        float* tmax = nullptr;
        if (spread_mfma(p, b)) {
            float* smax = pf_spread_smax(p, b, partial);
            if (!smax) return pf_fail(PF_ERR_ARG, "spread needs its workspace");
            tmax = reinterpret_cast<float*>(reinterpret_cast<char*>(smax) + al256((size_t)p->mtot * nvch * sizeof(float)));
            if (!smax_ready) {
                const int64_t threads = p->mtot * nvch * 16;
                hipLaunchKernelGGL(k_smax, dim3((unsigned)((threads + TB - 1) / TB)), dim3(TB), 0, s, p->mtot, b.C, ys,
                                   smax);
                PF_HIP(hipGetLastError());
            }
            const int64_t waves = (int64_t)I.nsitems * nvch;
            hipLaunchKernelGGL(k_tile_max, dim3((unsigned)((waves * 64 + TB - 1) / TB)), dim3(TB), 0, s, I.nsitems,
                               nvch, I.sitems, L.slist, smax, tmax);
            PF_HIP(hipGetLastError());
        }
        rc.fn(dim3((unsigned)nb), s, g, B, L, I, p, ys, grid, partial, tmax);
        PF_HIP(hipGetLastError());
        if (I.nritems > 0) {
            hipLaunchKernelGGL(k_spread_reduce, dim3((unsigned)((int64_t)I.nritems * nvch * nsub)), dim3(TB), 0, s, g,
                               B, VC, nsub, I.ritems, partial, grid);
            PF_HIP(hipGetLastError());
        }
        return PF_OK;
    }
    hipLaunchKernelGGL(k_spread, dim3((unsigned)nb), dim3(TB), lds, s, g, B, VC, I.sitems, L.slist, p->taprec,
                       p->perm, ys, grid, partial);
    PF_HIP(hipGetLastError());
    if (I.nritems > 0) {
        hipLaunchKernelGGL(k_spread_reduce, dim3((unsigned)((int64_t)I.nritems * nvch * nsub)), dim3(TB), 0, s, g, B,
                           VC, nsub, I.ritems, partial, grid);
        PF_HIP(hipGetLastError());
    }
    return PF_OK;
}

// Kernel attributes the operator launches rely on, set once per plan at
// creation (on the plan's device) rather than on the call path.
int pf_ops_prepare() {
    PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_spread), hipFuncAttributeMaxDynamicSharedMemorySize,
                               160 * 1024));
    return PF_OK;
}

static dim3 permute_grid(const pf_plan* p, const pf_batch& b) {
    return dim3((unsigned)((p->mtot + PSB - 1) / PSB), (unsigned)((b.C + PSC - 1) / PSC));
}

static dim3 permute_grid4(const pf_plan* p, const pf_batch& b) {
    return dim3((unsigned)((p->mtot + QSB - 1) / QSB), (unsigned)((b.C + QSC - 1) / QSC));
}

// the 16-byte permutation kernels need pairs of values on both sides
static bool permute_vec(const pf_plan* p, const pf_batch& b, const void* a, const void* c) {
    return b.C % 2 == 0 && p->mtot % 2 == 0 && ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(c)) & 15) == 0;
}

int pf_launch_sort(const pf_plan* p, const pf_batch& b, const float2* y, float2* ys, const float* weight,
                   hipStream_t s, float* smax, bool* smax_written) {
    if ((p->mtot + PSB - 1) / PSB > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many samples");
    // the maxima come with the 16-byte sort of whole 32-value chunks only
    if (!permute_vec(p, b, y, ys) || b.C % QSC != 0) smax = nullptr;
    if (smax_written) *smax_written = smax != nullptr;
    if (permute_vec(p, b, y, ys))
        hipLaunchKernelGGL(k_sort4, permute_grid4(p, b), dim3(TB), 0, s, p->mtot, b.C, p->iperm, y, weight, ys,
                           smax);
    else
        hipLaunchKernelGGL(k_sort, permute_grid(p, b), dim3(TB), 0, s, p->mtot, b.C, p->iperm, y, weight, ys);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

// the expanded sort's bounds: values per sample (its divisions are exact for
// n * d < 2^32 with n < 32 * V) and samples per block
constexpr int PF_EXPAND_VMAX = 1024;
constexpr int PF_EXPAND_QSH = 5;  // 32 samples per block

bool pf_spread_expand(const pf_plan* p, const pf_batch& b, pf_batch* bx) {
    if (!PF_SPREAD_EXPAND || (!b.b0t && !b.basis) || b.Cb != 1 || b.R * b.C > PF_EXPAND_VMAX) return false;
    pf_batch e = b;
    e.C = e.Cv = b.R * b.C;
    e.Cb = 1;
    e.R = e.L = e.K = 1;
    e.nt = 1;
    e.basis = e.b0t = e.b0s = nullptr;
    const RegChoice rc = spread_reg_choice(p, e);
    if (!rc.fn || !spread_dma(p, e, rc)) return false;
    if (bx) *bx = e;
    return true;
}

static uint64_t magic_div(int d) { return ((1ull << 32) + (uint64_t)d - 1) / (uint64_t)d; }

int pf_launch_sort_expand(const pf_plan* p, const pf_batch& b, const float2* y, float2* ys, const float* weight,
                          hipStream_t s) {
    const int V = b.R * b.C;
    if (V > PF_EXPAND_VMAX) return pf_fail(PF_ERR_UNSUPPORTED, "too many values per sample to expand");
    // samples per block: values, coefficients and rows within 32 KB of LDS
    int qsh = PF_EXPAND_QSH;
    auto lds = [&](int q) { return (size_t)q * ((b.C + 1 + b.R) * sizeof(float2) + sizeof(int)); };
    while (qsh > 0 && lds(1 << qsh) > 32768) --qsh;
    const int64_t nb = (p->mtot + (1 << qsh) - 1) >> qsh;
    if (nb > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many samples");
    if (nb < 1) return PF_OK;
    const bool pair = b.C % 2 == 0;
    const uint64_t m_out = magic_div(pair ? V / 2 : V), m_c = magic_div(b.C), m_r = magic_div(b.R);
    if (pair)
        hipLaunchKernelGGL(k_sort_x<true>, dim3((unsigned)nb), dim3(TB), lds(1 << qsh), s, geo(p), bparams(b), p->mtot,
                           qsh, m_out, m_c, m_r, p->iperm, p->taprec, y, weight, ys);
    else
        hipLaunchKernelGGL(k_sort_x<false>, dim3((unsigned)nb), dim3(TB), lds(1 << qsh), s, geo(p), bparams(b), p->mtot,
                           qsh, m_out, m_c, m_r, p->iperm, p->taprec, y, weight, ys);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_unsort(const pf_plan* p, const pf_batch& b, const float2* ys, float2* y, const float* weight,
                     float factor, hipStream_t s) {
    if ((p->mtot + PSB - 1) / PSB > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many samples");
    if (permute_vec(p, b, ys, y))
        hipLaunchKernelGGL(k_unsort4, permute_grid4(p, b), dim3(TB), 0, s, p->mtot, b.C, p->iperm, ys, weight, factor,
                           y);
    else
        hipLaunchKernelGGL(k_unsort, permute_grid(p, b), dim3(TB), 0, s, p->mtot, b.C, p->iperm, ys, weight, factor,
                           y);
    PF_HIP(hipGetLastError());
    return PF_OK;
}
