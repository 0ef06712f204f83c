// Coordinate gradient of the NUFFT: Kaiser-Bessel-derivative gather.
//
// For y_j = sum_r coef_r(j) sum_taps prod_a phi_a(tap_a - c_j,a) * grid_r[tap]
// (the forward operator before its 1/scale and norm factors), this kernel
// returns per sorted sample s and axis a
//
//     d[s, a] = Re( sum_c conj(h[s, c]) * dy[s, c] / dc_a ),
//
// where dphi/dc is the analytic derivative of the same Abramowitz & Stegun
// polynomials the tap weights come from, so the result is the exact derivative
// of the discrete operator.  With h = weight * (upstream gradient) this is the
// coordinate gradient of nufft; with h = weight * (adjoint input) and the grid
// built from the upstream image gradient, that of nufft_adj.  It replaces the
// incomplete _nufft_backward of src/pyfourier/_nufft/_nufft.py:175-209
// (reference defect A.14: no -2 pi i / N factor, invalid reshape).
#include <algorithm>

#include "pf_gcol.h"
#include "pf_internal.h"

namespace {

constexpr int TB = 256;

struct GradParams {
    int os[3];
    int width[3];
    float beta[3];
    int woff[3];
    int rec;
    int64_t gtot;
    int64_t npts;
    int64_t mtot;
    int C, Cv, Cb, R, K, nfg, nt;
    const float2* basis;
    const float2* b0t;
    int ndim;
    int accumulate;
    float scale[3];  // per stored axis (z, y, x)
};

// write one sample's three axis values at its original index, coordinate order
__device__ __forceinline__ void write_dcoord(const GradParams& P, const int32_t* perm, int64_t s, const float* g,
                                             float* dout) {
    const int64_t j = perm[s];
    for (int a = 3 - P.ndim; a < 3; ++a) {
        float* o = dout + j * P.ndim + (2 - a);
        const float v = P.scale[a] * g[a];
        *o = P.accumulate ? *o + v : v;
    }
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// d phi / d u for phi(u) = I0_AS(beta * sqrt(1 - u^2)), |u| <= 1 (0 outside)
__device__ double kb_deriv(double u, double beta) {
    if (fabs(u) > 1.0) return 0.0;
    const double v = beta * sqrt(1.0 - u * u);
    const double t = v / 3.75;
    if (v < 3.75) {
        // P = 1 + sum_k a_k t^(2k);  (dP/dv)/v = sum_k 2k a_k t^(2k-2) / 3.75^2
        const double t2 = t * t;
        const double q = 2.0 * 3.5156229 + t2 * (4.0 * 3.0899424 + t2 * (6.0 * 1.2067492 +
                         t2 * (8.0 * 0.2659732 + t2 * (10.0 * 0.0360768 + t2 * 12.0 * 0.0045813))));
        return -beta * beta * u * q / (3.75 * 3.75);
    }
    // Q = v^(-1/2) e^v p(s), s = 3.75 / v
    const double s = 1.0 / t;
    const double p = 0.39894228 + s * (0.01328592 + s * (0.00225319 + s * (-0.00157565 + s * (0.00916281 +
                     s * (-0.02057706 + s * (0.02635537 + s * (-0.01647633 + s * 0.00392377)))))));
    const double dp = 0.01328592 + s * (2 * 0.00225319 + s * (3 * -0.00157565 + s * (4 * 0.00916281 +
                      s * (5 * -0.02057706 + s * (6 * 0.02635537 + s * (7 * -0.01647633 + s * 8 * 0.00392377))))));
    const double e = exp(v) / sqrt(v);
    const double dq = e * (p * (1.0 - 0.5 / v) - dp * 3.75 / (v * v));
    return dq * (-beta * beta * u / v);
}

__global__ void __launch_bounds__(TB) k_gather_dcoord(GradParams P, const float* __restrict__ taprec,
                                                      const float* __restrict__ cos_sorted,
                                                      const int32_t* __restrict__ perm,
                                                      const float2* __restrict__ grid,
                                                      const float2* __restrict__ h, float* __restrict__ dout) {
    __shared__ float red[TB * 3];
    const int LS = P.C;  // one lane per (sample, batch item)
    const int SPB = TB / LS;
    const int64_t s = (int64_t)blockIdx.x * SPB + threadIdx.x / LS;
    const int lane = threadIdx.x % LS;
    const bool active = threadIdx.x < SPB * LS && s < P.mtot;
    float part[3] = {0.f, 0.f, 0.f};
    if (active) {
        const int c = lane;
        const int cb = c / P.Cv, cv = c % P.Cv;
        const float* rec = taprec + s * P.rec;
        const int* ri = reinterpret_cast<const int*>(rec);
        const int f = ri[3];
        const int fg = P.nfg > 1 ? f : 0;
        const int64_t tread = P.b0t ? ((int64_t)perm[s] % P.npts) % P.nt : 0;
        int idx[3][PF_WMAX];
        float w[3][PF_WMAX], dw[3][PF_WMAX];
        for (int a = 0; a < 3; ++a) {
            const int W = P.width[a];
            const double cd = (double)cos_sorted[s * 3 + a];
            const double half = 0.5 * W;
            const double x0 = W == 1 ? 0.0 : ceil(cd - half);
            for (int k = 0; k < W; ++k) {
                int v = ri[a] + k;
                idx[a][k] = v >= P.os[a] ? v - P.os[a] : v;
                w[a][k] = rec[P.woff[a] + k];
                // d phi(u)/d c with u = (x0 + k - c) / (W/2)
                dw[a][k] = W == 1 ? 0.f : (float)(-kb_deriv((x0 + k - cd) / half, (double)P.beta[a]) / half);
            }
        }
        const int RC = P.R * P.Cv;
        const float2* slab = grid + ((int64_t)fg * P.Cb + cb) * P.gtot * RC + cv;
        float2 D[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        for (int r = 0; r < P.R; ++r) {
            float2 acc[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
            const float2* sr = slab + r * P.Cv;
            for (int i = 0; i < P.width[0]; ++i)
                for (int j = 0; j < P.width[1]; ++j) {
                    const int64_t row = ((int64_t)idx[0][i] * P.os[1] + idx[1][j]) * P.os[2];
                    const float wzy = w[0][i] * w[1][j];
                    const float dz = dw[0][i] * w[1][j];
                    const float dy = w[0][i] * dw[1][j];
                    for (int k = 0; k < P.width[2]; ++k) {
                        const float2 g = sr[(row + idx[2][k]) * RC];
                        const float az = dz * w[2][k], ay = dy * w[2][k], ax = wzy * dw[2][k];
                        acc[0].x += az * g.x;
                        acc[0].y += az * g.y;
                        acc[1].x += ay * g.x;
                        acc[1].y += ay * g.y;
                        acc[2].x += ax * g.x;
                        acc[2].y += ax * g.y;
                    }
                }
            float2 cf = make_float2(1.f, 0.f);
            if (P.basis || P.b0t) {
                const int l = r / P.K, kk = r % P.K;
                if (P.b0t) cf = P.b0t[(int64_t)l * P.nt + tread];
                if (P.basis) cf = cmul(cf, P.basis[(int64_t)f * P.K + kk]);
            }
            for (int a = 0; a < 3; ++a) {
                const float2 m = cmul(acc[a], cf);
                D[a].x += m.x;
                D[a].y += m.y;
            }
        }
        const float2 hv = h[s * P.C + c];
        for (int a = 0; a < 3; ++a) part[a] = hv.x * D[a].x + hv.y * D[a].y;  // Re(conj(h) D)
    }
    for (int a = 0; a < 3; ++a) red[threadIdx.x * 3 + a] = part[a];
    __syncthreads();
    if (active && lane == 0) {
        float sum[3] = {0.f, 0.f, 0.f};
        for (int q = 0; q < LS; ++q)
            for (int a = 0; a < 3; ++a) sum[a] += red[(threadIdx.x + q) * 3 + a];
        write_dcoord(P, perm, s, sum, dout);
    }
}

// Same result with compile-time y/x widths: the lanes of a sample (CPL batch
// values each) first compute the sample's 3 x W derivative weights together
// into LDS, then every lane walks the taps with the y/x loops unrolled (all
// loads of a z row in flight) and 32-bit in-slab offsets.
// DEEP (3-D plans): the compiler may plan for 2 waves per SIMD and keep more
// tap loads in flight per wave, as in the forward gather's 3-D instances
template <int CPL, int WY, int WX, bool DEEP = false>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(1, DEEP ? 2 : 8)))
k_gather_dcoord_w(GradParams P, const float* __restrict__ taprec,
                                                        const float* __restrict__ cos_sorted,
                                                        const int32_t* __restrict__ perm,
                                                        const float2* __restrict__ grid,
                                                        const float2* __restrict__ h, float* __restrict__ dout) {
    __shared__ float red[TB * 3];
    __shared__ float sdw[TB][3 * PF_WMAX];  // indexed by the sample's slot in the block
    const int LS = P.C / CPL;
    const int SPB = TB / LS;
    const int slot = threadIdx.x / LS;
    const int64_t s = (int64_t)blockIdx.x * SPB + slot;
    const int lane = threadIdx.x % LS;
    const bool active = threadIdx.x < SPB * LS && s < P.mtot;
    const float* rec = taprec + (active ? s : 0) * P.rec;
    const int* ri = reinterpret_cast<const int*>(rec);
    if (active) {
        const int nw = P.width[0] + P.width[1] + P.width[2];
        for (int q = lane; q < nw; q += LS) {
            const int a = q < P.width[0] ? 0 : (q < P.width[0] + P.width[1] ? 1 : 2);
            const int k = q - (a == 0 ? 0 : (a == 1 ? P.width[0] : P.width[0] + P.width[1]));
            const int W = P.width[a];
            float d = 0.f;
            if (W > 1) {
                const double cd = (double)cos_sorted[s * 3 + a];
                const double half = 0.5 * W;
                const double x0 = ceil(cd - half);
                d = (float)(-kb_deriv((x0 + k - cd) / half, (double)P.beta[a]) / half);
            }
            sdw[slot][a * PF_WMAX + k] = d;
        }
    }
    __syncthreads();
    float part[3] = {0.f, 0.f, 0.f};
    if (active) {
        const int c0 = lane * CPL;
        const int cb = c0 / P.Cv, cv = c0 % P.Cv;
        const int f = ri[3];
        const int fg = P.nfg > 1 ? f : 0;
        const int64_t tread = P.b0t ? ((int64_t)perm[s] % P.npts) % P.nt : 0;
        const int RC = P.R * P.Cv;
        const float* dw = sdw[slot];
        int yo[WY], xo[WX];
        float wy[WY], wx[WX], dy[WY], dx[WX];
#pragma unroll
        for (int k = 0; k < WY; ++k) {
            int v = ri[1] + k;
            v -= v >= P.os[1] ? P.os[1] : 0;
            yo[k] = v * P.os[2] * RC;
            wy[k] = rec[P.woff[1] + k];
            dy[k] = dw[PF_WMAX + k];
        }
#pragma unroll
        for (int k = 0; k < WX; ++k) {
            int v = ri[2] + k;
            v -= v >= P.os[2] ? P.os[2] : 0;
            xo[k] = v * RC;
            wx[k] = rec[P.woff[2] + k];
            dx[k] = dw[2 * PF_WMAX + k];
        }
        const int zstride = P.os[1] * P.os[2] * RC;
        const float2* slab = grid + ((int64_t)fg * P.Cb + cb) * P.gtot * RC + cv;
        float2 D[3][CPL];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < CPL; ++q) D[a][q] = make_float2(0.f, 0.f);
        for (int r = 0; r < P.R; ++r) {
            float2 acc[3][CPL];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int q = 0; q < CPL; ++q) acc[a][q] = make_float2(0.f, 0.f);
            const float2* sr = slab + r * P.Cv;
            for (int i = 0; i < P.width[0]; ++i) {
                int z = ri[0] + i;
                z -= z >= P.os[0] ? P.os[0] : 0;
                const float wz = rec[P.woff[0] + i], dz = dw[i];
                const float2* pz = sr + z * zstride;
#pragma unroll
                for (int b = 0; b < WY; ++b) {
                    const float2* py = pz + yo[b];
                    // the row's x-weighted and x-derivative-weighted sums,
                    // folded into the three derivatives once per row
                    float2 s0[CPL], s1[CPL];
#pragma unroll
                    for (int q = 0; q < CPL; ++q) s0[q] = s1[q] = make_float2(0.f, 0.f);
#pragma unroll
                    for (int c = 0; c < WX; ++c) {
                        float2 g[CPL];
                        if constexpr (CPL == 2) {
                            const float4 v = *reinterpret_cast<const float4*>(py + xo[c]);
                            g[0] = make_float2(v.x, v.y);
                            g[CPL - 1] = make_float2(v.z, v.w);
                        } else {
                            g[0] = py[xo[c]];
                        }
#pragma unroll
                        for (int q = 0; q < CPL; ++q) {
                            s0[q].x += wx[c] * g[q].x;
                            s0[q].y += wx[c] * g[q].y;
                            s1[q].x += dx[c] * g[q].x;
                            s1[q].y += dx[c] * g[q].y;
                        }
                    }
                    const float wzy = wz * wy[b], dzy = dz * wy[b], wzdy = wz * dy[b];
#pragma unroll
                    for (int q = 0; q < CPL; ++q) {
                        acc[0][q].x += dzy * s0[q].x;
                        acc[0][q].y += dzy * s0[q].y;
                        acc[1][q].x += wzdy * s0[q].x;
                        acc[1][q].y += wzdy * s0[q].y;
                        acc[2][q].x += wzy * s1[q].x;
                        acc[2][q].y += wzy * s1[q].y;
                    }
                }
            }
            float2 cf = make_float2(1.f, 0.f);
            if (P.basis || P.b0t) {
                const int l = r / P.K, kk = r % P.K;
                if (P.b0t) cf = P.b0t[(int64_t)l * P.nt + tread];
                if (P.basis) cf = cmul(cf, P.basis[(int64_t)f * P.K + kk]);
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const float2 m = cmul(acc[a][q], cf);
                    D[a][q].x += m.x;
                    D[a][q].y += m.y;
                }
        }
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            const float2 hv = h[s * P.C + c0 + q];
#pragma unroll
            for (int a = 0; a < 3; ++a) part[a] += hv.x * D[a][q].x + hv.y * D[a][q].y;  // Re(conj(h) D)
        }
    }
    for (int a = 0; a < 3; ++a) red[threadIdx.x * 3 + a] = part[a];
    __syncthreads();
    if (active && lane == 0) {
        float sum[3] = {0.f, 0.f, 0.f};
        for (int q = 0; q < LS; ++q)
            for (int a = 0; a < 3; ++a) sum[a] += red[(threadIdx.x + q) * 3 + a];
        write_dcoord(P, perm, s, sum, dout);
    }
}

using GradLaunch = void (*)(dim3, hipStream_t, const GradParams&, const pf_plan*, const float2*, const float2*, float*);

template <int CPL, int WY, int WX, bool DEEP = false>
void launch_dcoord_w(dim3 grid, hipStream_t s, const GradParams& P, const pf_plan* p, const float2* g,
                     const float2* h, float* out) {
    hipLaunchKernelGGL((k_gather_dcoord_w<CPL, WY, WX, DEEP>), grid, dim3(TB), 0, s, P, p->taprec, p->cos_sorted,
                       p->perm, g, h, out);
}

// [cpl - 1][wy == 1 ? 0 : 1][wx - 1]
const GradLaunch DCOORD_W[2][2][PF_WMAX] = {
    {{&launch_dcoord_w<1, 1, 1>, &launch_dcoord_w<1, 1, 2>, &launch_dcoord_w<1, 1, 3>, &launch_dcoord_w<1, 1, 4>,
      &launch_dcoord_w<1, 1, 5>, &launch_dcoord_w<1, 1, 6>, &launch_dcoord_w<1, 1, 7>, &launch_dcoord_w<1, 1, 8>},
     {nullptr, &launch_dcoord_w<1, 2, 2>, &launch_dcoord_w<1, 3, 3>, &launch_dcoord_w<1, 4, 4>,
      &launch_dcoord_w<1, 5, 5>, &launch_dcoord_w<1, 6, 6>, &launch_dcoord_w<1, 7, 7>, &launch_dcoord_w<1, 8, 8>}},
    {{&launch_dcoord_w<2, 1, 1>, &launch_dcoord_w<2, 1, 2>, &launch_dcoord_w<2, 1, 3>, &launch_dcoord_w<2, 1, 4>,
      &launch_dcoord_w<2, 1, 5>, &launch_dcoord_w<2, 1, 6>, &launch_dcoord_w<2, 1, 7>, &launch_dcoord_w<2, 1, 8>},
     {nullptr, &launch_dcoord_w<2, 2, 2>, &launch_dcoord_w<2, 3, 3>, &launch_dcoord_w<2, 4, 4>,
      &launch_dcoord_w<2, 5, 5>, &launch_dcoord_w<2, 6, 6>, &launch_dcoord_w<2, 7, 7>, &launch_dcoord_w<2, 8, 8>}},
};

// square y/x widths of 3-D plans: [cpl - 1][w - 1]
const GradLaunch DCOORD_W_DEEP[2][PF_WMAX] = {
    {nullptr, &launch_dcoord_w<1, 2, 2, true>, &launch_dcoord_w<1, 3, 3, true>, &launch_dcoord_w<1, 4, 4, true>,
     &launch_dcoord_w<1, 5, 5, true>, &launch_dcoord_w<1, 6, 6, true>, &launch_dcoord_w<1, 7, 7, true>,
     &launch_dcoord_w<1, 8, 8, true>},
    {nullptr, &launch_dcoord_w<2, 2, 2, true>, &launch_dcoord_w<2, 3, 3, true>, &launch_dcoord_w<2, 4, 4, true>,
     &launch_dcoord_w<2, 5, 5, true>, &launch_dcoord_w<2, 6, 6, true>, &launch_dcoord_w<2, 7, 7, true>,
     &launch_dcoord_w<2, 8, 8, true>},
};


// LDS-staged derivative gather for the batches the forward gather stages (3-D,
// one width W = 3..8 on all axes, plain values, coil-pair-major grid: Cv = 2,
// Cb = C / 2).  Same work items, lane order and window staging as the forward
// LDS gather: one block per (frame, 16 x 16 column, DZ z planes) window, one
// lane per sample, the window of each coil pair copied to LDS and read from
// there.  Per (z, y) row a lane forms the x-weighted and x-derivative-weighted
// sums of its W taps and folds them into the three derivatives; Re(conj(h) D)
// accumulates over the coil pairs in registers, so each sample is written once
// and no cross-lane reduction is needed.
template <int W>
__global__ void __launch_bounds__(GCOL_TB, 2)
    k_gather_dcoord_col(GradParams P, const int4* __restrict__ items, const int32_t* __restrict__ gids,
                        const float* __restrict__ taprec, const float* __restrict__ cos_sorted,
                        const int32_t* __restrict__ perm, const float2* __restrict__ grid,
                        const float2* __restrict__ h, float* __restrict__ dout, int ncy, int ncx, int nzw) {
    using S = GColShape<W>;
    __shared__ float4 win[S::NPTS];
    const unsigned bid = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int4 it = items[bid];
    const int zw = it.x % nzw;
    const int cx = (it.x / nzw) % ncx;
    const int cy = (it.x / (nzw * ncx)) % ncy;
    const int f = it.x / (nzw * ncx * ncy);
    const int z0 = zw * S::DZ, y0 = cy * PF_GCOL, x0 = cx * PF_GCOL;
    const int ncp = P.C / 2;
    const float4* pairs = reinterpret_cast<const float4*>(grid) + (int64_t)(P.nfg > 1 ? f : 0) * P.Cb * P.gtot;
    const int tid = threadIdx.x, wave = tid >> 6;

    int goff[S::NK];
#pragma unroll
    for (int k = 0; k < S::NK; ++k) {
        const int p = k * GCOL_TB + tid;
        const int pl = p / S::PP, rem = p - pl * S::PP;
        const int py = rem / S::PX, px = rem - py * S::PX;
        const int z = (z0 + pl) % P.os[0], y = (y0 + py) % P.os[1], x = (x0 + px) % P.os[2];
        goff[k] = p < S::NPTS ? (z * P.os[1] + y) * P.os[2] + x : -1;
    }

    const int t0 = gids[it.y + tid];
    const bool act = t0 >= 0;
    const int t = act ? t0 : 0;
    int base = 0;
    float w[3][W], dw[3][W];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < W; ++k) w[a][k] = dw[a][k] = 0.f;
    if (act) {
        const float* rec = taprec + (int64_t)t * P.rec;
        const int* ri = reinterpret_cast<const int*>(rec);
        base = ((ri[0] - z0) * S::PX + (ri[1] - y0)) * S::PX + (ri[2] - x0);
        constexpr double half = 0.5 * W;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double cd = (double)cos_sorted[(int64_t)t * 3 + a];
            const double x0d = ceil(cd - half);
#pragma unroll
            for (int k = 0; k < W; ++k) {
                w[a][k] = rec[P.woff[a] + k];
                dw[a][k] = (float)(-kb_deriv((x0d + k - cd) / half, (double)P.beta[a]) / half);
            }
        }
    }
    const float4* hp = reinterpret_cast<const float4*>(h + (int64_t)t * P.C);
    float part[3] = {0.f, 0.f, 0.f};

    for (int cp = 0; cp < ncp; ++cp) {
        const float4* pair = pairs + (int64_t)cp * P.gtot;
#pragma unroll
        for (int k = 0; k < S::NK; ++k)
            if (goff[k] >= 0) gcol_copy16(pair + goff[k], win + k * GCOL_TB + wave * 64);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (act) {
            float4 D[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) D[a] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int a = 0; a < W; ++a)
#pragma unroll
                for (int b = 0; b < W; ++b) {
                    const float4* row = win + base + a * S::PP + b * S::PX;
                    float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
#pragma unroll
                    for (int c = 0; c < W; ++c) {
                        const float4 v = row[c];
                        s0.x += w[2][c] * v.x;
                        s0.y += w[2][c] * v.y;
                        s0.z += w[2][c] * v.z;
                        s0.w += w[2][c] * v.w;
                        s1.x += dw[2][c] * v.x;
                        s1.y += dw[2][c] * v.y;
                        s1.z += dw[2][c] * v.z;
                        s1.w += dw[2][c] * v.w;
                    }
                    const float dzy = dw[0][a] * w[1][b], wzdy = w[0][a] * dw[1][b], wzy = w[0][a] * w[1][b];
                    D[0].x += dzy * s0.x;
                    D[0].y += dzy * s0.y;
                    D[0].z += dzy * s0.z;
                    D[0].w += dzy * s0.w;
                    D[1].x += wzdy * s0.x;
                    D[1].y += wzdy * s0.y;
                    D[1].z += wzdy * s0.z;
                    D[1].w += wzdy * s0.w;
                    D[2].x += wzy * s1.x;
                    D[2].y += wzy * s1.y;
                    D[2].z += wzy * s1.z;
                    D[2].w += wzy * s1.w;
                }
            const float4 hv = hp[cp];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                part[a] += hv.x * D[a].x + hv.y * D[a].y + hv.z * D[a].z + hv.w * D[a].w;  // Re(conj(h) D)
        }
        __syncthreads();  // the next pair's copies overwrite the window
    }
    if (act) write_dcoord(P, perm, t, part, dout);
}

template <int W>
void launch_dcoord_col(hipStream_t s, const GradParams& P, const pf_plan* p, const float2* g, const float2* h,
                       float* out) {
    hipLaunchKernelGGL((k_gather_dcoord_col<W>), dim3((unsigned)p->gl.nitems), dim3(GCOL_TB), 0, s, P, p->gl.items,
                       p->gl.ids, p->taprec, p->cos_sorted, p->perm, g, h, out, p->gl.ncol[0], p->gl.ncol[1],
                       p->gl.nzw);
}

using DcoordColLaunch = void (*)(hipStream_t, const GradParams&, const pf_plan*, const float2*, const float2*, float*);

// [W - 3], W = 3 .. 8
const DcoordColLaunch DCOORD_COL[6] = {&launch_dcoord_col<3>, &launch_dcoord_col<4>, &launch_dcoord_col<5>,
                                       &launch_dcoord_col<6>, &launch_dcoord_col<7>, &launch_dcoord_col<8>};

}  // namespace

int pf_launch_gather_dcoord(const pf_plan* p, const pf_batch& b, const float2* grid, const float2* h, const float* scale,
                            bool accumulate, float* out, hipStream_t s) {
    if (b.C > TB) return pf_fail(PF_ERR_UNSUPPORTED, "coordinate gradient supports at most 256 batch items");
    GradParams P;
    for (int a = 0; a < 3; ++a) {
        P.os[a] = p->os[a];
        P.width[a] = p->width[a];
        P.beta[a] = p->beta[a];
        P.woff[a] = p->woff[a];
    }
    P.rec = p->rec;
    P.gtot = pf_gtot(p);
    P.npts = p->npts;
    P.mtot = p->mtot;
    P.C = b.C;
    P.Cv = b.Cv;
    P.Cb = b.Cb;
    P.R = b.R;
    P.K = b.K;
    P.nfg = b.nfg;
    P.nt = b.nt;
    P.basis = b.basis;
    P.b0t = b.b0t;
    P.ndim = p->ndim;
    P.accumulate = accumulate ? 1 : 0;
    for (int a = 0; a < 3; ++a) P.scale[a] = scale[a];
    if (b.Cv == 2 && pf_gather_col_layout(p, b)) {
        DCOORD_COL[p->width[2] - 3](s, P, p, grid, h, out);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    const int cpl = (b.Cv % 2 == 0 && b.C >= 2) ? 2 : 1;
    const int wy = p->width[1], wx = p->width[2];
    const bool idx32 = (int64_t)pf_gtot(p) * b.R * b.Cv < (1ll << 31);
    if (idx32 && (wy == 1 || wy == wx)) {
        const bool deep = p->width[0] > 1 && wy == wx && wy > 1;
        if (GradLaunch fn = deep ? DCOORD_W_DEEP[cpl - 1][wx - 1] : DCOORD_W[cpl - 1][wy == 1 ? 0 : 1][wx - 1]) {
            const int spb = TB / (b.C / cpl);
            fn(dim3((unsigned)((p->mtot + spb - 1) / spb)), s, P, p, grid, h, out);
            PF_HIP(hipGetLastError());
            return PF_OK;
        }
    }
    const int SPB = TB / b.C;
    const int64_t nblk = (p->mtot + SPB - 1) / SPB;
    hipLaunchKernelGGL(k_gather_dcoord, dim3((unsigned)nblk), dim3(TB), 0, s, P, p->taprec, p->cos_sorted, p->perm,
                       grid, h, out);
    PF_HIP(hipGetLastError());
    return PF_OK;
}
