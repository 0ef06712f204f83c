// GRAPPA operator gridding (GROG) on gfx950.
//
// Replaces the per-sample coil-operator interpolation of
// src/pyfourier/_preinterp/_grog.py:22-249 (do_interpolation / _interp at
// :354-425) and the index / shift bookkeeping of :176-218.  For a sample at
// grid coordinate c (grid units, axis order x, y, z) the nearest Cartesian
// point is round(c); the shift d = round(c) - c (|d| <= 1/2) is quantised to
// nsteps levels, step_a = floor((nsteps - 1) * d_a) + nsteps / 2, and the
// sample's coil vector (a row vector, as the operators were trained,
// T = S G) is moved onto the grid point by
//
//     y = x @ Gx^(delta[step_x]) @ Gy^(delta[step_y]) [@ Gz^(delta[step_z])]
//
// with the fractional operator powers precomputed per axis and step (host,
// once per call).  The reference multiplies the product table
// G = Gx^dx Gy^dy Gz^dz; applying the three factors in turn is the same
// linear map and needs only 3 x nsteps matrices instead of nsteps^3, so a
// block stages one axis' table (nsteps x nc x nc) in LDS and every sample of
// its tile reads its own step's matrix there.  Samples stay in the caller's
// coil-major layout (B, nc, npts): loads and stores are unit-stride across
// the block's consecutive samples, and data cross HBM once each way.
#include <algorithm>
#include <cmath>

#include "pf_internal.h"

namespace {

constexpr int GT = 512;           // samples (threads) per block
constexpr int GJC = 16;           // output coils per inner chunk
constexpr int GLDS_MAX = 19200;   // LDS table budget in float2 (150 KB)

__device__ __forceinline__ float2 gmac(float2 acc, float2 a, float2 b) {
    return make_float2(acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x);
}

// per sample: packed step index (or -1 outside the grid), clamped grid index
// (x, y, z order) and the flattened index before clamping
__global__ void k_grog_index(const float* __restrict__ coord, int64_t npts, int ndim, int3 n_xyz, int nsteps,
                             int* __restrict__ lut, int* __restrict__ index, long long* __restrict__ flat) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npts) return;
    const int n[3] = {n_xyz.x, n_xyz.y, n_xyz.z};
    int packed = 0, mul = 1;
    bool inside = true;
    long long f = 0, unfold = 1;
    for (int a = 0; a < ndim; ++a) {
        const float c = coord[s * ndim + a];
        const float r = rintf(c);  // round half to even, like torch.round
        const float d = r - c;
        const float q = __fmul_rn((float)(nsteps - 1), d);
        const int step = (int)floorf(q) + nsteps / 2;
        packed += mul * min(max(step, 0), nsteps - 1);
        mul *= nsteps;
        const int idx = (int)r + n[a] / 2;
        f += unfold * (long long)idx;
        unfold *= n[a];
        int cl = idx;
        if (idx < 0) {
            inside = false;
            cl = 0;
        } else if (idx >= n[a]) {
            inside = false;
            cl = n[a] - 1;
        }
        index[s * ndim + a] = cl;
    }
    lut[s] = inside ? packed : -1;
    flat[s] = f;
}

// y = x @ T_a[step_a] for a = 0 .. ndim-1 (x, y, z); (B, nc, npts) layout.
// TABLE_LDS: the axis table of the block's operator slice sits in LDS;
// otherwise the lanes read it from global memory (large nc or nsteps).
template <int NCMAX, bool TABLE_LDS>
__global__ void __launch_bounds__(GT) k_grog_apply(const float2* __restrict__ in, float2* __restrict__ out,
                                                   int64_t npts, int nc, int ndim, int nsteps,
                                                   const float2* __restrict__ tables, int nob,
                                                   const int* __restrict__ lut) {
    extern __shared__ float4 tab4[];
    float2* tab = reinterpret_cast<float2*>(tab4);
    const int64_t ntile = (npts + GT - 1) / GT;
    const int64_t b = blockIdx.x / ntile;
    const int64_t s = (blockIdx.x % ntile) * GT + threadIdx.x;
    const int ob = nob > 1 ? (int)b : 0;
    // LDS rows padded with zeros to a multiple of GJC columns (no guards in
    // the inner loop, 16-byte reads); +2 per matrix spreads the steps' equal
    // entries over different banks
    const int ncp = (nc + GJC - 1) / GJC * GJC;
    const int mstride = nc * ncp + 2;
    const bool live = s < npts;
    const int code = live ? lut[s] : -1;
    const int64_t col = b * nc * npts + s;  // element (b, 0, s); coil i at + i * npts
    int rem = code < 0 ? 0 : code;
    for (int a = 0; a < ndim; ++a) {
        const int step = rem % nsteps;
        rem /= nsteps;
        const float2* T_g = tables + (((int64_t)a * nsteps) * nob + ob) * nc * nc;
        if (TABLE_LDS) {
            __syncthreads();  // previous axis done with the table
            for (int q = threadIdx.x; q < nsteps * nc * ncp; q += GT) {
                const int st = q / (nc * ncp), e = q % (nc * ncp);
                const int i = e / ncp, j = e % ncp;
                tab[st * mstride + e] = j < nc ? T_g[((int64_t)st * nob) * nc * nc + i * nc + j] : make_float2(0.f, 0.f);
            }
            __syncthreads();
        }
        if (!live) continue;
        const float2* src = a == 0 ? in : out;
        float2 x[NCMAX];
#pragma unroll
        for (int i = 0; i < NCMAX; ++i) x[i] = i < nc ? src[col + (int64_t)i * npts] : make_float2(0.f, 0.f);
        const float2* m = TABLE_LDS ? tab + step * mstride : T_g + (int64_t)step * nob * nc * nc;
        for (int j0 = 0; j0 < nc; j0 += GJC) {
            float2 y[GJC];
#pragma unroll
            for (int j = 0; j < GJC; ++j) y[j] = make_float2(0.f, 0.f);
#pragma unroll
            for (int i = 0; i < NCMAX; ++i) {
                if (i < nc) {
                    if (TABLE_LDS) {
                        const float4* row = reinterpret_cast<const float4*>(m + i * ncp + j0);
                        float4 r[GJC / 2];
#pragma unroll
                        for (int q = 0; q < GJC / 2; ++q) r[q] = row[q];
#pragma unroll
                        for (int q = 0; q < GJC / 2; ++q) {
                            y[2 * q] = gmac(y[2 * q], x[i], make_float2(r[q].x, r[q].y));
                            y[2 * q + 1] = gmac(y[2 * q + 1], x[i], make_float2(r[q].z, r[q].w));
                        }
                    } else {
                        const float2* row = m + i * nc + j0;
#pragma unroll
                        for (int j = 0; j < GJC; ++j)
                            if (j0 + j < nc) y[j] = gmac(y[j], x[i], row[j]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < GJC; ++j)
                if (j0 + j < nc) out[col + (int64_t)(j0 + j) * npts] = code < 0 ? make_float2(0.f, 0.f) : y[j];
        }
    }
}

template <int NCMAX>
int launch_apply(const float2* in, float2* out, int B, int nc, int64_t npts, int ndim, int nsteps,
                 const float2* tables, int nob, const int* lut, hipStream_t s) {
    const int64_t ntile = (npts + GT - 1) / GT;
    const int64_t nblk = (int64_t)B * ntile;
    if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many GROG samples");
    const int64_t tab = (int64_t)nsteps * (nc * ((nc + GJC - 1) / GJC * GJC) + 2);
    if (tab <= GLDS_MAX) {
        const size_t lds = (size_t)tab * sizeof(float2);
        if (lds > 64 * 1024) {
            static bool raised = false;
            if (!raised) {
                if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_grog_apply<NCMAX, true>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
                    return pf_fail(PF_ERR_HIP, "could not raise the GROG kernel's LDS limit");
                raised = true;
            }
        }
        hipLaunchKernelGGL((k_grog_apply<NCMAX, true>), dim3((unsigned)nblk), dim3(GT), lds, s, in, out, npts, nc,
                           ndim, nsteps, tables, nob, lut);
    } else {
        hipLaunchKernelGGL((k_grog_apply<NCMAX, false>), dim3((unsigned)nblk), dim3(GT), 0, s, in, out, npts, nc,
                           ndim, nsteps, tables, nob, lut);
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

}  // namespace

extern "C" {

int pf_grog_index(const float* coord, int ndim, const int* shape, long long npts, int nsteps, int* lut, int* index,
                  long long* flat, void* stream) {
    if (!coord || !shape || !lut || !index || !flat) return pf_fail(PF_ERR_ARG, "null argument");
    if (ndim < 2 || ndim > 3) return pf_fail(PF_ERR_ARG, "GROG needs 2 or 3 spatial dims");
    if (npts < 0) return pf_fail(PF_ERR_ARG, "negative sample count");
    if (nsteps < 1 || (long long)nsteps * nsteps * (ndim == 3 ? nsteps : 1) > (1ll << 30))
        return pf_fail(PF_ERR_ARG, "nsteps out of range");
    if (npts == 0) return PF_OK;
    int3 n = make_int3(1, 1, 1);
    // shape is (z, y, x) like every grid argument; the kernel works in (x, y, z)
    n.x = shape[ndim - 1];
    n.y = shape[ndim - 2];
    if (ndim == 3) n.z = shape[0];
    const int64_t nb = (npts + 255) / 256;
    hipLaunchKernelGGL(k_grog_index, dim3((unsigned)nb), dim3(256), 0, static_cast<hipStream_t>(stream), coord,
                       (int64_t)npts, ndim, n, nsteps, lut, index, flat);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_grog_apply(const void* in, void* out, int batch, int ncoil, long long npts, int ndim, int nsteps,
                  const void* tables, int table_batch, const int* lut, void* stream) {
    if (!in || !out || !tables || !lut) return pf_fail(PF_ERR_ARG, "null argument");
    if (in == out) return pf_fail(PF_ERR_ARG, "GROG input and output must not alias");
    if (batch < 1 || ncoil < 1 || npts < 0) return pf_fail(PF_ERR_ARG, "bad sizes");
    if (ndim < 2 || ndim > 3) return pf_fail(PF_ERR_ARG, "GROG needs 2 or 3 spatial dims");
    if (nsteps < 1) return pf_fail(PF_ERR_ARG, "nsteps must be >= 1");
    if (table_batch != 1 && table_batch != batch)
        return pf_fail(PF_ERR_ARG, "operator tables must be shared (1) or one per batch entry");
    if (ncoil > 64) return pf_fail(PF_ERR_UNSUPPORTED, "GROG supports up to 64 coils");
    if (npts == 0) return PF_OK;
    const float2* i2 = static_cast<const float2*>(in);
    float2* o2 = static_cast<float2*>(out);
    const float2* t2 = static_cast<const float2*>(tables);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (ncoil <= 8) return launch_apply<8>(i2, o2, batch, ncoil, npts, ndim, nsteps, t2, table_batch, lut, s);
    if (ncoil <= 16) return launch_apply<16>(i2, o2, batch, ncoil, npts, ndim, nsteps, t2, table_batch, lut, s);
    if (ncoil <= 32) return launch_apply<32>(i2, o2, batch, ncoil, npts, ndim, nsteps, t2, table_batch, lut, s);
    return launch_apply<64>(i2, o2, batch, ncoil, npts, ndim, nsteps, t2, table_batch, lut, s);
}

}  // extern "C"
