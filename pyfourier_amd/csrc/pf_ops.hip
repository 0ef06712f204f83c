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

#include "pf_internal.h"

namespace {

constexpr int TB = 256;

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
};

BatchParams bparams(const pf_batch& b) {
    BatchParams q{b.C, b.Cv, b.Cb, b.R, b.L, b.K, b.nfg, b.nimg, b.nt, b.basis, b.b0t, b.b0s};
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

// -------------------------------------------------------------------------
// spread: one block per (tile work item, coil block, value chunk).  Waves
// own disjoint value subsets of the LDS tile; inside a wave all lanes work on
// the same sample, on distinct (tap line, value) pairs, so no two lanes ever
// add to the same LDS word in one instruction and per-wave program order
// fixes the summation order: run-to-run bit-identical.
// -------------------------------------------------------------------------
constexpr int SPREAD_LDS_BYTES = 64 * 1024;

__global__ void __launch_bounds__(TB) k_spread(GeoParams g, BatchParams B, int VC,
                                               const int4* __restrict__ items, int nitems,
                                               const int32_t* __restrict__ slist,
                                               const float* __restrict__ taprec,
                                               const int32_t* __restrict__ perm,
                                               const float2* __restrict__ ys,
                                               float2* __restrict__ grid, float2* __restrict__ partial) {
    extern __shared__ float lds[];
    const int nvch = (B.R * B.Cv + VC - 1) / VC;
    unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    int4 it = items[bid];
    const int cb = blockIdx.y / nvch;
    const int vch = blockIdx.y % nvch;
    const int RC = B.R * B.Cv;
    const int vbase = vch * VC;
    const int vcnt = min(VC, RC - vbase);
    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tx = (int)(tile % g.ntiles[2]);
    const int ty = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int ox = tx * g.tile[2], oy = ty * g.tile[1], oz = tz * g.tile[0];
    const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
    const int ex = min(Tx, g.os[2] - ox), ey = min(Ty, g.os[1] - oy), ez = min(Tz, g.os[0] - oz);
    const int tpts = Tx * Ty * Tz;

    for (int i = threadIdx.x; i < tpts * VC * 2; i += blockDim.x) lds[i] = 0.f;
    __syncthreads();

    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int nw = blockDim.x / 64;
    // values owned by this wave: [vw0, vw1)
    const int per = (vcnt + nw - 1) / nw;
    const int vw0 = min(vcnt, wave * per), vw1 = min(vcnt, vw0 + per);
    const int VW = vw1 - vw0;
    const int Wz = g.width[0], Wy = g.width[1], Wx = g.width[2];
    const int lines = Wz * Wy;
    const int npairs = lines * VW;

    if (VW > 0) {
        for (int e = it.y; e < it.z; ++e) {
            const int64_t s = slist[e];
            const float* rec = taprec + s * g.rec;
            const int* ri = reinterpret_cast<const int*>(rec);
            const int f = ri[3];
            int64_t tread = 0;
            if (B.b0t) tread = ((int64_t)perm[s] % g.npts) % B.nt;
            for (int q = lane; q < npairs; q += 64) {
                const int line = q / VW;
                const int v = vw0 + q % VW;
                const int a = line / Wy, b = line % Wy;
                int gz = ri[0] + a;
                if (gz >= g.os[0]) gz -= g.os[0];
                int gy = ri[1] + b;
                if (gy >= g.os[1]) gy -= g.os[1];
                const int lz = gz - oz, ly = gy - oy;
                if (lz < 0 || lz >= ez || ly < 0 || ly >= ey) continue;
                const int vg = vbase + v;
                const int r = vg / B.Cv, cv = vg % B.Cv;
                float2 val = ys[s * B.C + cb * B.Cv + cv];
                if (B.basis || B.b0t) val = cmulc(val, sample_coef(B, r, f, tread));
                const float wzy = rec[g.woff[0] + a] * rec[g.woff[1] + b];
                float* rowp = lds + ((int64_t)(lz * Ty + ly) * Tx) * VC * 2 + v * 2;
                for (int c = 0; c < Wx; ++c) {
                    int gx = ri[2] + c;
                    if (gx >= g.os[2]) gx -= g.os[2];
                    const int lx = gx - ox;
                    if (lx < 0 || lx >= ex) continue;
                    const float w = wzy * rec[g.woff[2] + c];
                    float* d = rowp + lx * VC * 2;
                    atomicAdd(d, w * val.x);
                    atomicAdd(d + 1, w * val.y);
                }
            }
        }
    }
    __syncthreads();

    if (it.w < 0) {
        float2* slab = grid + ((int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + cb) * g.gtot * RC + vbase;
        for (int i = threadIdx.x; i < tpts * vcnt; i += blockDim.x) {
            int v = i % vcnt;
            int pt = i / vcnt;
            int lx = pt % Tx, ly = (pt / Tx) % Ty, lz = pt / (Tx * Ty);
            if (lx >= ex || ly >= ey || lz >= ez) continue;
            int64_t gpt = ((int64_t)(oz + lz) * g.os[1] + (oy + ly)) * g.os[2] + (ox + lx);
            const float* src = lds + ((int64_t)pt * VC + v) * 2;
            slab[gpt * RC + v] = make_float2(src[0], src[1]);
        }
    } else {
        float2* dst = partial + (((int64_t)it.w * B.Cb + cb) * nvch + vch) * (int64_t)tpts * VC;
        for (int i = threadIdx.x; i < tpts * VC; i += blockDim.x)
            dst[i] = make_float2(lds[2 * i], lds[2 * i + 1]);
    }
}

// Sum the partial tiles of a heavy tile in slot order and write it out.
__global__ void __launch_bounds__(TB) k_spread_reduce(GeoParams g, BatchParams B, int VC,
                                                      const int4* __restrict__ ritems,
                                                      const float2* __restrict__ partial,
                                                      float2* __restrict__ grid) {
    const int nvch = (B.R * B.Cv + VC - 1) / VC;
    int4 it = ritems[blockIdx.x];
    const int cb = blockIdx.y / nvch;
    const int vch = blockIdx.y % nvch;
    const int RC = B.R * B.Cv;
    const int vbase = vch * VC;
    const int vcnt = min(VC, RC - vbase);
    const int64_t key = it.x;
    const int fg_frame = (int)(key / g.ntile_tot);
    const int64_t tile = key % g.ntile_tot;
    const int tx = (int)(tile % g.ntiles[2]);
    const int ty = (int)((tile / g.ntiles[2]) % g.ntiles[1]);
    const int tz = (int)(tile / ((int64_t)g.ntiles[2] * g.ntiles[1]));
    const int ox = tx * g.tile[2], oy = ty * g.tile[1], oz = tz * g.tile[0];
    const int Tx = g.tile[2], Ty = g.tile[1], Tz = g.tile[0];
    const int ex = min(Tx, g.os[2] - ox), ey = min(Ty, g.os[1] - oy), ez = min(Tz, g.os[0] - oz);
    const int tpts = Tx * Ty * Tz;
    float2* slab = grid + ((int64_t)(B.nfg > 1 ? fg_frame : 0) * B.Cb + cb) * g.gtot * RC + vbase;
    for (int i = threadIdx.x; i < tpts * vcnt; i += blockDim.x) {
        int v = i % vcnt;
        int pt = i / vcnt;
        int lx = pt % Tx, ly = (pt / Tx) % Ty, lz = pt / (Tx * Ty);
        if (lx >= ex || ly >= ey || lz >= ez) continue;
        float2 acc = make_float2(0.f, 0.f);
        for (int sl = 0; sl < it.z; ++sl) {
            const float2* src = partial + (((int64_t)(it.y + sl) * B.Cb + cb) * nvch + vch) * (int64_t)tpts * VC;
            float2 v2 = src[(int64_t)pt * VC + v];
            acc.x += v2.x;
            acc.y += v2.y;
        }
        int64_t gpt = ((int64_t)(oz + lz) * g.os[1] + (oy + ly)) * g.os[2] + (ox + lx);
        slab[gpt * RC + v] = acc;
    }
}

// user (C, F, npts) -> sorted (Mtot, C), times weight
__global__ void k_sort(int64_t mtot, int C, const int32_t* __restrict__ iperm, const float2* __restrict__ y,
                       const float* __restrict__ w, float2* __restrict__ ys) {
    const int64_t total = mtot * C;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = e % mtot;
        int c = (int)(e / mtot);
        float2 v = y[e];
        if (w) {
            float a = w[j];
            v.x *= a;
            v.y *= a;
        }
        ys[(int64_t)iperm[j] * C + c] = v;
    }
}

// sorted (Mtot, C) -> user (C, F, npts), times factor and weight
__global__ void k_unsort(int64_t mtot, int C, const int32_t* __restrict__ iperm, const float2* __restrict__ ys,
                         const float* __restrict__ w, float factor, float2* __restrict__ y) {
    const int64_t total = mtot * C;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = e % mtot;
        int c = (int)(e / mtot);
        float2 v = ys[(int64_t)iperm[j] * C + c];
        float a = factor * (w ? w[j] : 1.0f);
        y[e] = make_float2(v.x * a, v.y * a);
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

int pf_launch_gather(const pf_plan* p, const pf_batch& b, const float2* grid, float2* ys, hipStream_t s) {
    GeoParams g = geo(p);
    BatchParams B = bparams(b);
    if (b.Cv % 2 == 0) {
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

int pf_spread_coil_chunk(const pf_plan* p, const pf_batch& b) {
    int tpts = p->tile[0] * p->tile[1] * p->tile[2];
    int vc = SPREAD_LDS_BYTES / (tpts * 8);
    if (vc < 1) vc = 1;
    return std::min(vc, b.R * b.Cv);
}

size_t pf_spread_partial_bytes(const pf_plan* p, const pf_batch& b) {
    int VC = pf_spread_coil_chunk(p, b);
    int nvch = (b.R * b.Cv + VC - 1) / VC;
    size_t tpts = (size_t)p->tile[0] * p->tile[1] * p->tile[2];
    return (size_t)pf_slist_for(p, b).nslots * b.Cb * nvch * tpts * VC * sizeof(float2);
}

int pf_launch_spread(const pf_plan* p, const pf_batch& b, const float2* ys, float2* grid, float2* partial,
                     hipStream_t s) {
    GeoParams g = geo(p);
    BatchParams B = bparams(b);
    int VC = pf_spread_coil_chunk(p, b);
    int nvch = (b.R * b.Cv + VC - 1) / VC;
    int tpts = p->tile[0] * p->tile[1] * p->tile[2];
    size_t lds = (size_t)tpts * VC * 8;
    const pf_slist& L = pf_slist_for(p, b);
    if (L.nslots > 0 && !partial) return pf_fail(PF_ERR_ARG, "spread needs a partial-tile workspace");
    if ((int64_t)b.Cb * nvch > 65535) return pf_fail(PF_ERR_UNSUPPORTED, "too many coil chunks");
    hipLaunchKernelGGL(k_spread, dim3(L.nsitems, b.Cb * nvch), dim3(TB), lds, s, g, B, VC, L.sitems, L.nsitems,
                       L.slist, p->taprec, p->perm, ys, grid, partial);
    PF_HIP(hipGetLastError());
    if (L.nritems > 0) {
        hipLaunchKernelGGL(k_spread_reduce, dim3(L.nritems, b.Cb * nvch), dim3(TB), 0, s, g, B, VC, L.ritems,
                           partial, grid);
        PF_HIP(hipGetLastError());
    }
    return PF_OK;
}

int pf_launch_sort(const pf_plan* p, const pf_batch& b, const float2* y, float2* ys, const float* weight,
                   hipStream_t s) {
    int64_t total = p->mtot * b.C;
    hipLaunchKernelGGL(k_sort, dim3(grid_stride_blocks(total)), dim3(TB), 0, s, p->mtot, b.C, p->iperm, y,
                       weight, ys);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_unsort(const pf_plan* p, const pf_batch& b, const float2* ys, float2* y, const float* weight,
                     float factor, hipStream_t s) {
    int64_t total = p->mtot * b.C;
    hipLaunchKernelGGL(k_unsort, dim3(grid_stride_blocks(total)), dim3(TB), 0, s, p->mtot, b.C, p->iperm, ys,
                       weight, factor, y);
    PF_HIP(hipGetLastError());
    return PF_OK;
}
