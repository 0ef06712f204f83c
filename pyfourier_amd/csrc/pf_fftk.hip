// Pruned, fused Cartesian FFT stage of the NUFFT (gfx950).
//
// Forward  (replaces _apodize + _resize + centred fftn, src/pyfourier/_nufft/_nufft.py:234-241):
//   X pass   image rows (coil-major user layout) -> apodise (x B0 map) ->
//            place at the stored positions (p - N/2) mod G (pad + ifftshift)
//            -> FFT along x -> T1 (nfg, Cb, Nz, Ny, Gx, R*Cv)
//   Y pass   T1 (Ny compact rows) -> zero-padded FFT along y -> T2 / grid
//   Z pass   T2 (Nz compact planes) -> zero-padded FFT along z -> grid
// Only rows that hold image data are transformed: in 3-D at sigma=2 the X pass
// works on 1/4 and the Y pass on 1/2 of the grid's lines.
//
// Adjoint  (replaces centred ifftn + _resize + _apodize, src/pyfourier/_nufft/_nufft.py:276-282):
//   the mirror image, each pass keeping only the N outputs the crop needs,
//   the X pass writing the deapodised image (sum over B0 segments with the
//   conjugate maps) in the user layout.
//
// Lines of G = 16 Q points (Q in {5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20,
// 21, 24, 25, 28, 30, 32, 36, 40}: 80-640) and of G = 8 Q points (Q in {5,
// 7, 9, 15, 21, 25, 35, 45}: 40-360) run register four-step kernels: each
// thread holds Q values of a line, takes a small DFT in registers, twiddles,
// and one LDS transpose hands the line to the second small DFT over the 16
// or 8 threads (k_fftcol256/512, k_fftcolq, k_fftx256, k_fftx512_*,
// k_fftxq).  Other sizes: each block
// transforms B columns of one line family held in LDS as [G][B+1] complex,
// with an in-place mixed-radix (8, 5, 4, 3, 2) Stockham FFT (Govindaraju et
// al., "High performance discrete Fourier transforms on graphics
// processors", SC 2008).  Twiddles come from a per-axis table of
// exp(-2 pi i k / G) computed in double on the host.  The inverse uses
// conj(FFT(conj(x))).
#include <algorithm>
#include <cmath>
#include <vector>

#include "pf_internal.h"

namespace {

#ifndef PF_FFT_THREADS
#define PF_FFT_THREADS 256
#endif
#ifndef PF_FFT_GBMAX
#define PF_FFT_GBMAX 4096
#endif
constexpr int FT = PF_FFT_THREADS;  // threads per block
// minimum waves per SIMD the FFT kernels are compiled for: 4 caps them at
// 128 VGPRs, so four 256-thread blocks (the LDS limit at G*B = 4096) fit a CU
#ifndef PF_FFT_WPE
#define PF_FFT_WPE 4
#endif
constexpr int GBMAX = PF_FFT_GBMAX;  // complex values per block (G * B)

struct AxisFFT {
    int G;
    int N;
    int nrad;
    int rad[16];
    uint64_t nsm[16];  // division magic of Ns before each pass
    const float2* tw;  // (G,) exp(-2 pi i k / G)
};

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_mulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 c_negi(float2 a) { return make_float2(a.y, -a.x); }  // -i * a
__device__ __forceinline__ float2 c_posi(float2 a) { return make_float2(-a.y, a.x); }  // +i * a
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// forward DFTs (kernel exp(-2 pi i r k / R)), in place on v[0..R)
__device__ __forceinline__ void dft2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = c_add(a, b);
    v[1] = c_sub(a, b);
}
__device__ __forceinline__ void dft3(float2* v) {
    const float s = 0.86602540378443864676f;
    float2 t = c_add(v[1], v[2]);
    float2 d = c_sub(v[1], v[2]);
    float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    float2 r = c_negi(c_scale(d, s));
    v[0] = c_add(v[0], t);
    v[1] = c_add(m, r);
    v[2] = c_sub(m, r);
}
__device__ __forceinline__ void dft4(float2* v) {
    float2 a = c_add(v[0], v[2]), b = c_sub(v[0], v[2]);
    float2 c = c_add(v[1], v[3]), d = c_negi(c_sub(v[1], v[3]));
    v[0] = c_add(a, c);
    v[2] = c_sub(a, c);
    v[1] = c_add(b, d);
    v[3] = c_sub(b, d);
}
__device__ __forceinline__ void dft5(float2* v) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 p14 = c_add(v[1], v[4]), m14 = c_sub(v[1], v[4]);
    float2 p23 = c_add(v[2], v[3]), m23 = c_sub(v[2], v[3]);
    float2 x0 = v[0];
    float2 a1 = make_float2(x0.x + c1 * p14.x + c2 * p23.x, x0.y + c1 * p14.y + c2 * p23.y);
    float2 a2 = make_float2(x0.x + c2 * p14.x + c1 * p23.x, x0.y + c2 * p14.y + c1 * p23.y);
    float2 b1 = make_float2(s1 * m14.x + s2 * m23.x, s1 * m14.y + s2 * m23.y);
    float2 b2 = make_float2(s2 * m14.x - s1 * m23.x, s2 * m14.y - s1 * m23.y);
    v[0] = make_float2(x0.x + p14.x + p23.x, x0.y + p14.y + p23.y);
    v[1] = c_add(a1, c_negi(b1));
    v[4] = c_add(a1, c_posi(b1));
    v[2] = c_add(a2, c_negi(b2));
    v[3] = c_add(a2, c_posi(b2));
}
__device__ __forceinline__ void dft8(float2* v) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e);
    dft4(o);
    float2 o1 = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));   // W8^1 = (1-i)/sqrt2
    float2 o2 = c_negi(o[2]);                                               // W8^2 = -i
    float2 o3 = make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));  // W8^3 = (-1-i)/sqrt2
    v[0] = c_add(e[0], o[0]);
    v[4] = c_sub(e[0], o[0]);
    v[1] = c_add(e[1], o1);
    v[5] = c_sub(e[1], o1);
    v[2] = c_add(e[2], o2);
    v[6] = c_sub(e[2], o2);
    v[3] = c_add(e[3], o3);
    v[7] = c_sub(e[3], o3);
}

// forward 7-point DFT: with a_j = x_j + x_(7-j), b_j = x_j - x_(7-j) (j = 1..3),
// X_k = x_0 + sum_j a_j cos(2 pi j k / 7) - i sum_j b_j sin(2 pi j k / 7)
// and X_(7-k) the same with the sine terms negated
__device__ __forceinline__ void dft7(float2* v) {
    const float c1 = 0.62348980185873359f, c2 = -0.22252093395631434f, c3 = -0.90096886790241903f;
    const float s1 = 0.7818314824680298f, s2 = 0.97492791218182362f, s3 = 0.43388373911755823f;
    const float2 x0 = v[0];
    const float2 a1 = c_add(v[1], v[6]), b1 = c_sub(v[1], v[6]);
    const float2 a2 = c_add(v[2], v[5]), b2 = c_sub(v[2], v[5]);
    const float2 a3 = c_add(v[3], v[4]), b3 = c_sub(v[3], v[4]);
    // k = 1, 2, 3: cosine parts r_k and sine parts q_k (x the pair differences)
    const float2 r1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x, x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y);
    const float2 r2 = make_float2(x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x, x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y);
    const float2 r3 = make_float2(x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x, x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y);
    const float2 q1 = make_float2(s1 * b1.x + s2 * b2.x + s3 * b3.x, s1 * b1.y + s2 * b2.y + s3 * b3.y);
    const float2 q2 = make_float2(s2 * b1.x - s3 * b2.x - s1 * b3.x, s2 * b1.y - s3 * b2.y - s1 * b3.y);
    const float2 q3 = make_float2(s3 * b1.x - s1 * b2.x + s2 * b3.x, s3 * b1.y - s1 * b2.y + s2 * b3.y);
    v[0] = make_float2(x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y);
    v[1] = c_add(r1, c_negi(q1));
    v[6] = c_add(r1, c_posi(q1));
    v[2] = c_add(r2, c_negi(q2));
    v[5] = c_add(r2, c_posi(q2));
    v[3] = c_add(r3, c_negi(q3));
    v[4] = c_add(r3, c_posi(q3));
}

template <int R>
__device__ __forceinline__ void dftR(float2* v) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 5) dft5(v);
    else if constexpr (R == 7) dft7(v);
    else dft8(v);
}

// floor(q / d) for q < 8192, d <= 4096 with m = ceil(2^32 / d) (exact in that range)
__device__ __forceinline__ int mdiv(int q, uint64_t m) { return (int)(((uint64_t)(unsigned)q * m) >> 32); }
__host__ __device__ inline uint64_t magic_of(int d) { return ((1ull << 32) + (uint64_t)d - 1) / (uint64_t)d; }

// One in-place Stockham pass of radix R over B columns (stride LD) of length G.
// bm, nm: division magics of B and Ns.
template <int R>
__device__ void stockham_pass(float2* buf, int LD, int B, uint64_t bm, int G, int Ns, uint64_t nm,
                              const float2* tw) {
    constexpr int CMAX = (GBMAX + R * FT - 1) / (R * FT);
    const int nb = G / R;
    const int total = nb * B;
    const int tstep = G / (Ns * R);
    float2 v[CMAX][R];
#pragma unroll
    for (int c = 0; c < CMAX; ++c) {
        const int q = threadIdx.x + c * FT;
        if (q < total) {
            const int j = mdiv(q, bm), b = q - j * B;
            const int k = j - mdiv(j, nm) * Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float2 x = buf[(j + r * nb) * LD + b];
                v[c][r] = r == 0 ? x : c_mul(x, tw[r * k * tstep]);
            }
            dftR<R>(v[c]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CMAX; ++c) {
        const int q = threadIdx.x + c * FT;
        if (q < total) {
            const int j = mdiv(q, bm), b = q - j * B;
            const int jn = mdiv(j, nm), k = j - jn * Ns;
            const int d = jn * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[(d + r * Ns) * LD + b] = v[c][r];
        }
    }
    __syncthreads();
}

// Full FFT of the LDS block.  Caller has synchronised after filling it.
__device__ void lds_fft(float2* buf, int LD, int B, const AxisFFT& A, const float2* tw) {
    const uint64_t bm = magic_of(B);
    int Ns = 1;
    for (int p = 0; p < A.nrad; ++p) {
        const uint64_t nm = A.nsm[p];
        switch (A.rad[p]) {
            case 8: stockham_pass<8>(buf, LD, B, bm, A.G, Ns, nm, tw); break;
            case 5: stockham_pass<5>(buf, LD, B, bm, A.G, Ns, nm, tw); break;
            case 4: stockham_pass<4>(buf, LD, B, bm, A.G, Ns, nm, tw); break;
            case 3: stockham_pass<3>(buf, LD, B, bm, A.G, Ns, nm, tw); break;
            default: stockham_pass<2>(buf, LD, B, bm, A.G, Ns, nm, tw); break;
        }
        Ns *= A.rad[p];
    }
}

__device__ __forceinline__ int stored_pos(int p, int N, int G) {
    int n = p - N / 2;
    return n < 0 ? n + G : n;
}

struct Geo {
    int N[3], G[3];   // (z, y, x)
    int64_t ntot;     // Nz*Ny*Nx
};

struct Batch {
    int C, Cv, Cb, R, L, K, nfg, nimg;
    const float2* b0s;  // (L, ntot)
};

// ---------------------------------------------------------------------------
// X pass.  A block takes YG compact image rows (fg, cb, z, y) and a chunk of
// KC (coefficient, coil) pairs kc = k*Cv + cv with all L B0 segments, so the
// adjoint can sum the segments in-block.  Column j of a row holds value
// v = l*(K*Cv) + kc of the grid vector (v = r*Cv + cv, r = l*K + k).
// ---------------------------------------------------------------------------
struct XBlk {
    int YG;      // rows per block
    int KC;      // (k, cv) pairs per block
    int nkc;     // chunks over K*Cv
};

__device__ __forceinline__ void x_row_decode(const Batch& bt, int64_t row, int64_t nrows, int& fg, int& cb,
                                             int64_t& zy) {
    zy = row % nrows;
    int64_t gc = row / nrows;
    cb = (int)(gc % bt.Cb);
    fg = (int)(gc / bt.Cb);
}

__global__ void __launch_bounds__(FT, PF_FFT_WPE) k_fftx_fwd(Geo g, Batch bt, AxisFFT A, const float* __restrict__ apod,
                                                 const float2* __restrict__ img, float2* __restrict__ out, XBlk xb) {
    extern __shared__ float2 sh[];
    const int RC = bt.R * bt.Cv;
    const int KCv = bt.K * bt.Cv;
    const int VC = bt.L * xb.KC;
    const int B = xb.YG * VC, LD = B + 1;
    const int G = A.G, N = A.N;
    float2* buf = sh;
    float2* tw = sh + G * LD;
    const int64_t nrows = (int64_t)g.N[0] * g.N[1];
    const int64_t rows_tot = (int64_t)bt.nfg * bt.Cb * nrows;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t row0 = (blk / xb.nkc) * xb.YG;
    const int kc0 = (int)(blk % xb.nkc) * xb.KC;
    for (int i = threadIdx.x; i < G; i += FT) tw[i] = A.tw[i];
    for (int i = threadIdx.x; i < G * LD; i += FT) buf[i] = make_float2(0.f, 0.f);
    __syncthreads();
    const float* apz = apod;
    const float* apy = apod + g.N[0];
    const float* apx = apod + g.N[0] + g.N[1];
    for (int q = threadIdx.x; q < B * N; q += FT) {
        const int p = q % N;
        const int col = q / N;
        const int yy = col / VC, j = col % VC;
        const int l = j / xb.KC, kc = kc0 + j % xb.KC;
        const int64_t row = row0 + yy;
        if (row >= rows_tot || kc >= KCv) continue;
        int fg, cb;
        int64_t zy;
        x_row_decode(bt, row, nrows, fg, cb, zy);
        const int k = kc / bt.Cv, cv = kc % bt.Cv;
        const int z = (int)(zy / g.N[1]), y = (int)(zy % g.N[1]);
        const int ii = bt.nfg > 1 ? fg : (bt.nimg > 1 ? k : 0);
        const int c = cb * bt.Cv + cv;
        const int64_t pix = zy * N + p;
        float2 x = img[((int64_t)c * bt.nimg + ii) * g.ntot + pix];
        x = c_scale(x, apz[z] * apy[y] * apx[p]);
        if (bt.b0s) x = c_mul(x, bt.b0s[(int64_t)l * g.ntot + pix]);
        buf[stored_pos(p, N, G) * LD + col] = x;
    }
    __syncthreads();
    lds_fft(buf, LD, B, A, tw);
    for (int q = threadIdx.x; q < xb.YG * G * VC; q += FT) {
        const int j = q % VC;
        const int n = (q / VC) % G;
        const int yy = q / (VC * G);
        const int l = j / xb.KC, kc = kc0 + j % xb.KC;
        const int64_t row = row0 + yy;
        if (row >= rows_tot || kc >= KCv) continue;
        out[(row * G + n) * RC + l * KCv + kc] = buf[n * LD + yy * VC + j];
    }
}

__global__ void __launch_bounds__(FT, PF_FFT_WPE) k_fftx_adj(Geo g, Batch bt, AxisFFT A, const float* __restrict__ apod,
                                                 const float2* __restrict__ in, float2* __restrict__ img, XBlk xb,
                                                 float factor) {
    extern __shared__ float2 sh[];
    const int RC = bt.R * bt.Cv;
    const int KCv = bt.K * bt.Cv;
    const int VC = bt.L * xb.KC;
    const int B = xb.YG * VC, LD = B + 1;
    const int G = A.G, N = A.N;
    float2* buf = sh;
    float2* tw = sh + G * LD;
    const int64_t nrows = (int64_t)g.N[0] * g.N[1];
    const int64_t rows_tot = (int64_t)bt.nfg * bt.Cb * nrows;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t row0 = (blk / xb.nkc) * xb.YG;
    const int kc0 = (int)(blk % xb.nkc) * xb.KC;
    for (int i = threadIdx.x; i < G; i += FT) tw[i] = A.tw[i];
    for (int q = threadIdx.x; q < xb.YG * G * VC; q += FT) {
        const int j = q % VC;
        const int n = (q / VC) % G;
        const int yy = q / (VC * G);
        const int l = j / xb.KC, kc = kc0 + j % xb.KC;
        const int64_t row = row0 + yy;
        float2 x = make_float2(0.f, 0.f);
        if (row < rows_tot && kc < KCv) x = in[(row * G + n) * RC + l * KCv + kc];
        buf[n * LD + yy * VC + j] = make_float2(x.x, -x.y);  // conj: inverse via forward FFT
    }
    __syncthreads();
    lds_fft(buf, LD, B, A, tw);
    const float* apz = apod;
    const float* apy = apod + g.N[0];
    const float* apx = apod + g.N[0] + g.N[1];
    for (int q = threadIdx.x; q < xb.YG * xb.KC * N; q += FT) {
        const int p = q % N;
        const int kcl = (q / N) % xb.KC;
        const int yy = q / (N * xb.KC);
        const int kc = kc0 + kcl;
        const int64_t row = row0 + yy;
        if (row >= rows_tot || kc >= KCv) continue;
        int fg, cb;
        int64_t zy;
        x_row_decode(bt, row, nrows, fg, cb, zy);
        const int k = kc / bt.Cv, cv = kc % bt.Cv;
        const int z = (int)(zy / g.N[1]), y = (int)(zy % g.N[1]);
        const int64_t pix = zy * N + p;
        const int n = stored_pos(p, N, G);
        float2 acc = make_float2(0.f, 0.f);
        for (int l = 0; l < bt.L; ++l) {
            float2 x = buf[n * LD + yy * VC + l * xb.KC + kcl];
            x = make_float2(x.x, -x.y);
            if (bt.b0s) x = c_mulc(x, bt.b0s[(int64_t)l * g.ntot + pix]);
            acc = c_add(acc, x);
        }
        const int ii = bt.nfg > 1 ? fg : (bt.nimg > 1 ? k : 0);
        const int c = cb * bt.Cv + cv;
        img[((int64_t)c * bt.nimg + ii) * g.ntot + pix] = c_scale(acc, factor * apz[z] * apy[y] * apx[p]);
    }
}

// ---------------------------------------------------------------------------
// Column pass along a strided axis.  Buffer view: (O, rows, I) with I the
// contiguous inner extent; a block takes B inner columns of one outer index.
// fwd: N compact input rows -> G output rows (zero-padded FFT)
// adj: G input rows -> inverse FFT -> N compact output rows
// ---------------------------------------------------------------------------
template <bool FWD>
__global__ void __launch_bounds__(FT, PF_FFT_WPE) k_fftcol(AxisFFT A, int64_t I, int B, int bshift,
                                               const float2* __restrict__ in, float2* __restrict__ out) {
    extern __shared__ float2 sh[];
    const int G = A.G, N = A.N;
    const int LD = B + 1;
    float2* buf = sh;
    float2* tw = sh + G * LD;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i0 = (blk % nchunk) * B;
    const int bw = (int)std::min<int64_t>(B, I - i0);
    for (int i = threadIdx.x; i < G; i += FT) tw[i] = A.tw[i];
    if (FWD) {
        // rows that receive no image data (stored positions N - N/2 .. G - N/2 - 1)
        const int z0 = N - N / 2;
        for (int i = threadIdx.x; i < (G - N) * LD; i += FT) buf[z0 * LD + i] = make_float2(0.f, 0.f);
        const float2* src = in + o * (int64_t)N * I + i0;
        for (int q = threadIdx.x; q < (N << bshift); q += FT) {
            const int b = q & (B - 1), p = q >> bshift;
            buf[stored_pos(p, N, G) * LD + b] = b < bw ? src[(int64_t)p * I + b] : make_float2(0.f, 0.f);
        }
    } else {
        const float2* src = in + o * (int64_t)G * I + i0;
        for (int q = threadIdx.x; q < (G << bshift); q += FT) {
            const int b = q & (B - 1), n = q >> bshift;
            const float2 x = b < bw ? src[(int64_t)n * I + b] : make_float2(0.f, 0.f);
            buf[n * LD + b] = make_float2(x.x, -x.y);
        }
    }
    __syncthreads();
    lds_fft(buf, LD, B, A, tw);
    if (FWD) {
        float2* dst = out + o * (int64_t)G * I + i0;
        for (int q = threadIdx.x; q < (G << bshift); q += FT) {
            const int b = q & (B - 1), n = q >> bshift;
            if (b < bw) dst[(int64_t)n * I + b] = buf[n * LD + b];
        }
    } else {
        float2* dst = out + o * (int64_t)N * I + i0;
        for (int q = threadIdx.x; q < (N << bshift); q += FT) {
            const int b = q & (B - 1), p = q >> bshift;
            if (b < bw) {
                const float2 x = buf[stored_pos(p, N, G) * LD + b];
                dst[(int64_t)p * I + b] = make_float2(x.x, -x.y);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Column pass for G = 256 as a four-step FFT in registers: element n = t + 16k
// of a column is held by thread t (k = 0..15); thread t takes the 16-point DFT
// over k, multiplies by exp(-2 pi i t m1 / 256), the block transposes through
// LDS once, and thread m1 takes the 16-point DFT over t, giving output
// m1 + 16 m2.  Two LDS accesses per value instead of two per radix pass, and
// zero rows of the padded input are not loaded.  Lane = (column b, t), b
// fastest, so global accesses are 16 consecutive columns of one row.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dft4v(float2& a, float2& b, float2& c, float2& d) {
    const float2 s0 = c_add(a, c), d0 = c_sub(a, c);
    const float2 s1 = c_add(b, d), d1 = c_negi(c_sub(b, d));
    a = c_add(s0, s1);
    c = c_sub(s0, s1);
    b = c_add(d0, d1);
    d = c_sub(d0, d1);
}

// forward 16-point DFT in place (kernel exp(-2 pi i n m / 16)), v[n] -> v[m]
__device__ __forceinline__ void dft16(float2* v) {
    // n = 4a + b: DFT4 over a for each b, twiddle W16^(b c), DFT4 over b
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;  // W16^1 = c1 - i s1
    const float h = 0.70710678118654752440f;                                  // W16^2 = h - i h
#pragma unroll
    for (int b = 0; b < 4; ++b) dft4v(v[b], v[4 + b], v[8 + b], v[12 + b]);
    // Y_b[c] sits at v[4c + b]; multiply by W16^(b c)
    v[4 * 1 + 1] = c_mul(v[5], make_float2(c1, -s1));
    v[4 * 1 + 2] = c_mul(v[6], make_float2(h, -h));
    v[4 * 1 + 3] = c_mul(v[7], make_float2(s1, -c1));
    v[4 * 2 + 1] = c_mul(v[9], make_float2(h, -h));
    v[4 * 2 + 2] = c_negi(v[10]);
    v[4 * 2 + 3] = c_mul(v[11], make_float2(-h, -h));
    v[4 * 3 + 1] = c_mul(v[13], make_float2(s1, -c1));
    v[4 * 3 + 2] = c_mul(v[14], make_float2(-h, -h));
    v[4 * 3 + 3] = c_mul(v[15], make_float2(-c1, s1));
    // DFT4 over b for each c: X[c + 4d]
    float2 o[16];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float2 a0 = v[4 * c], a1 = v[4 * c + 1], a2 = v[4 * c + 2], a3 = v[4 * c + 3];
        dft4v(a0, a1, a2, a3);
        o[c] = a0;
        o[c + 4] = a1;
        o[c + 8] = a2;
        o[c + 12] = a3;
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = o[m];
}

// Dead grid columns of a 3-D plan (no sample's footprint covers the (y, x)
// column) in the fused adjoint's inverse z pass (inner index
// i = (gy*Gx + gx)*RC + rc): the spread leaves their empty tiles unwritten,
// so a dead column is taken as zero without reading it (its zero output is
// still written).  mode 0: every column read.  Skipping dead columns in the
// y passes, or their stores in the forward, measured slower (per-row mask
// reads, rows written with holes) or no faster on the 3-D radial headline.
struct ColMask {
    const uint8_t* m;
    int mode, RC;
};

__device__ __forceinline__ bool cm_dead_col(const ColMask& cm, int64_t i) {
    return cm.mode == 1 && !cm.m[i / cm.RC];
}

// Columns per block of the 256-point column passes: a block reads and writes
// BC consecutive values of each row, so a larger BC gives longer contiguous
// runs (BC x 8 bytes) per row at the price of LDS (16 x (16 BC + 8) float2).
// Measured on the 3-D headline: forward passes fastest at 64, inverse at 32
// (16: pad+FFT 2.80 / IFFT+crop 2.47 ms; 32: 2.61 / 2.29; 64: 2.52 / 2.46).
constexpr int C256_BF = 64;                 // forward
constexpr int C256_BA = 32;                 // inverse
// columns per block of the Cartesian column passes (x-contiguous lines: 16
// measured fastest, masked 256^3 x 8 coils fwd+adj 3.12 ms vs 3.17-3.23 ms
// for 32 or 64)
constexpr int CC256_BF = 16;
constexpr int CC256_BA = 16;

template <bool FWD, int B>
__global__ void __launch_bounds__(16 * B, PF_FFT_WPE * 16 / B) k_fftcol256(int N, int64_t I,
                                                                         const float2* __restrict__ twg,
                                                                         const float2* __restrict__ in,
                                                                         float2* __restrict__ out, ColMask cm) {
    constexpr int G = 256;
    constexpr int C256_PLANE = 16 * B + 8;  // LDS plane stride (float2) of the transpose
    __shared__ float2 sh[16 * C256_PLANE];
    __shared__ float2 tw[G];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i0 = (blk % nchunk) * B;
    const bool bok = i0 + b < I;
    const bool zero = !bok || cm_dead_col(cm, i0 + b);  // no column, or one known to be zero
    if (threadIdx.x < G) tw[threadIdx.x] = twg[threadIdx.x];
    const int lo = N - N / 2, hi = G - N / 2;  // stored rows [lo, hi) carry no image data
    float2 v[16];
    if (FWD) {
        const float2* src = in + o * (int64_t)N * I + i0 + b;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int n = t + 16 * k;  // stored row
            const int p = n < lo ? n + N / 2 : n - G + N / 2;
            v[k] = (!zero && (n < lo || n >= hi)) ? src[(int64_t)p * I] : make_float2(0.f, 0.f);
        }
    } else {
        const float2* src = in + o * (int64_t)G * I + i0 + b;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const float2 x = !zero ? src[(int64_t)(t + 16 * k) * I] : make_float2(0.f, 0.f);
            v[k] = make_float2(x.x, -x.y);
        }
    }
    dft16(v);
    __syncthreads();  // twiddle table
#pragma unroll
    for (int m1 = 1; m1 < 16; ++m1) v[m1] = c_mul(v[m1], tw[t * m1]);
#pragma unroll
    for (int m1 = 0; m1 < 16; ++m1) sh[m1 * C256_PLANE + t * B + b] = v[m1];
    __syncthreads();
    // this thread now owns m1 = t and reads Y_t'[m1] for t' = 0..15
#pragma unroll
    for (int tp = 0; tp < 16; ++tp) v[tp] = sh[t * C256_PLANE + tp * B + b];
    dft16(v);  // v[m2] = X[t + 16 m2]
    if (!bok) return;
    if (FWD) {
        float2* dst = out + o * (int64_t)G * I + i0 + b;
#pragma unroll
        for (int m2 = 0; m2 < 16; ++m2) dst[(int64_t)(t + 16 * m2) * I] = v[m2];
    } else {
        float2* dst = out + o * (int64_t)N * I + i0 + b;
#pragma unroll
        for (int m2 = 0; m2 < 16; ++m2) {
            const int n = t + 16 * m2;  // stored output index -> cropped row p
            if (n < lo || n >= hi) {
                const int p = n < lo ? n + N / 2 : n - G + N / 2;
                dst[(int64_t)p * I] = make_float2(v[m2].x, -v[m2].y);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Column pass for G = 512 (256-point images at oversampling 2), the register
// four-step with 512 = 16 x 32: element n = t + 16k of a column is held by
// thread t (t = 0..15, k = 0..31), which takes the 32-point DFT over k (two
// 16-point DFTs and a radix-2 step) and multiplies by exp(-2 pi i t m1 / 512);
// the block transposes through LDS in two halves (m1 < 16, then m1 >= 16, so
// a half-size buffer keeps two blocks on a CU), and after each half thread t
// takes the 16-point DFT over t for m1 = t (+16), giving outputs m1 + 32 m2.
// Every thread works in both stages.  Lane = (column b, t), b fastest, B
// consecutive columns per block: each row access is one B x 8-byte run.
// ---------------------------------------------------------------------------
#ifndef PF_FFT512
#define PF_FFT512 1  // register four-step column passes for 512-point axes
#endif
#ifndef PF_FFTQ
#define PF_FFTQ 1  // register four-step column passes for 16 Q-point axes
#endif
#ifndef PF_FFT512X32
#define PF_FFT512X32 1  // 512-point inverse row passes with one segment as 16 threads x 32 values
#endif
constexpr int C512_B = 32;        // columns per block (512 threads)
constexpr int C512_BS = 16;       // k_fftcol512_t32: columns per block (512 threads)
constexpr int64_t C512_MIN = 1 << 18;  // columns from which k_fftcol512 runs (measured: 2^15 columns
                                       // 6% slower than k_fftcol512_t32, 2^18 even, 2^20-2^21 8-10% faster)

template <bool FWD, int B>
__global__ void __launch_bounds__(16 * B) k_fftcol512(int N, int64_t I, const float2* __restrict__ twg,
                                                    const float2* __restrict__ in, float2* __restrict__ out,
                                                    ColMask cm) {
    constexpr int G = 512;
    constexpr int PLANE = 16 * B;  // a 32-lane access covers 32 consecutive columns: no bank conflicts
    __shared__ float2 sh[16 * PLANE];
    __shared__ float2 tw[G];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i0 = (blk % nchunk) * B;
    const bool bok = i0 + b < I;
    const bool zero = !bok || cm_dead_col(cm, i0 + b);  // no column, or one known to be zero
    for (int k = threadIdx.x; k < G; k += 16 * B) tw[k] = twg[k];
    const int lo = N - N / 2, hi = G - N / 2;  // stored rows [lo, hi) carry no image data
    // v[s] = element k = 2s, v[16 + s] = element k = 2s + 1
    float2 v[32];
    if (FWD) {
        const float2* src = in + o * (int64_t)N * I + i0 + b;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int n = t + 16 * k;  // stored row
            const int p = n < lo ? n + N / 2 : n - G + N / 2;
            v[(k & 1) * 16 + (k >> 1)] =
                (!zero && (n < lo || n >= hi)) ? src[(int64_t)p * I] : make_float2(0.f, 0.f);
        }
    } else {
        const float2* src = in + o * (int64_t)G * I + i0 + b;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const float2 x = !zero ? src[(int64_t)(t + 16 * k) * I] : make_float2(0.f, 0.f);
            v[(k & 1) * 16 + (k >> 1)] = make_float2(x.x, -x.y);
        }
    }
    dft16(v);
    dft16(v + 16);
    __syncthreads();  // twiddle table
    // Y[m] = E[m] + W32^m O[m], Y[m + 16] = E[m] - W32^m O[m], times W512^(t m1)
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const float2 w = c_mul(v[16 + m], tw[16 * m]);
        const float2 e = v[m];
        v[m] = c_mul(c_add(e, w), tw[t * m]);
        v[16 + m] = c_mul(c_sub(e, w), tw[t * (m + 16)]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h) __syncthreads();  // the first half's reads are done
#pragma unroll
        for (int m = 0; m < 16; ++m) sh[m * PLANE + t * B + b] = v[16 * h + m];
        __syncthreads();
        float2 u[16];  // thread t: m1 = 16 h + t, u[t'] = Y_t'[m1]
#pragma unroll
        for (int tp = 0; tp < 16; ++tp) u[tp] = sh[t * PLANE + tp * B + b];
        dft16(u);  // u[m2] = X[m1 + 32 m2]
        if (bok) {
            const int m1 = 16 * h + t;
            if (FWD) {
                float2* dst = out + o * (int64_t)G * I + i0 + b;
#pragma unroll
                for (int m2 = 0; m2 < 16; ++m2) dst[(int64_t)(m1 + 32 * m2) * I] = u[m2];
            } else {
                float2* dst = out + o * (int64_t)N * I + i0 + b;
#pragma unroll
                for (int m2 = 0; m2 < 16; ++m2) {
                    const int n = m1 + 32 * m2;  // stored output index -> cropped row p
                    if (n < lo || n >= hi) {
                        const int p = n < lo ? n + N / 2 : n - G + N / 2;
                        dst[(int64_t)p * I] = make_float2(u[m2].x, -u[m2].y);
                    }
                }
            }
        }
    }
}

// The same transform with 512 = 32 x 16 for passes of fewer columns: 32
// threads per column hold 16 values each (stride 32) for a 16-point DFT,
// twiddle W512^(t m1), one LDS transpose, then threads m1 < 16 take the
// 32-point DFT over t as two 16-point DFTs and a radix-2 step (the upper
// half idles).  Fewer registers (100 / 140 against 195 / 172 VGPRs), so
// twice the threads per CU, which pays where a pass has few columns.
template <bool FWD, int B>
__global__ void __launch_bounds__(32 * B) k_fftcol512_t32(int N, int64_t I, const float2* __restrict__ twg,
                                                    const float2* __restrict__ in, float2* __restrict__ out,
                                                    ColMask cm) {
    constexpr int G = 512;
    constexpr int PLANE = 32 * B + 8;  // LDS plane stride (float2) of the transpose
    __shared__ float2 sh[16 * PLANE];
    __shared__ float2 tw[G];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i0 = (blk % nchunk) * B;
    const bool bok = i0 + b < I;
    const bool zero = !bok || cm_dead_col(cm, i0 + b);  // no column, or one known to be zero
    for (int k = threadIdx.x; k < G; k += 32 * B) tw[k] = twg[k];
    const int lo = N - N / 2, hi = G - N / 2;  // stored rows [lo, hi) carry no image data
    float2 v[32];
    if (FWD) {
        const float2* src = in + o * (int64_t)N * I + i0 + b;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int n = t + 32 * k;  // stored row
            const int p = n < lo ? n + N / 2 : n - G + N / 2;
            v[k] = (!zero && (n < lo || n >= hi)) ? src[(int64_t)p * I] : make_float2(0.f, 0.f);
        }
    } else {
        const float2* src = in + o * (int64_t)G * I + i0 + b;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const float2 x = !zero ? src[(int64_t)(t + 32 * k) * I] : make_float2(0.f, 0.f);
            v[k] = make_float2(x.x, -x.y);
        }
    }
    dft16(v);
    __syncthreads();  // twiddle table
#pragma unroll
    for (int m1 = 1; m1 < 16; ++m1) v[m1] = c_mul(v[m1], tw[t * m1]);
#pragma unroll
    for (int m1 = 0; m1 < 16; ++m1) sh[m1 * PLANE + t * B + b] = v[m1];
    __syncthreads();
    if (t >= 16 || !bok) return;
    // thread m1 = t: Y_t'[m1] for t' = 0..31, even t' into v[0..15], odd into v[16..31]
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        v[s] = sh[t * PLANE + (2 * s) * B + b];
        v[16 + s] = sh[t * PLANE + (2 * s + 1) * B + b];
    }
    dft16(v);
    dft16(v + 16);
    // X[t + 16 q] = E[q] + W32^q O[q], X[t + 16 (q + 16)] = E[q] - W32^q O[q]
    if (FWD) {
        float2* dst = out + o * (int64_t)G * I + i0 + b;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const float2 w = c_mul(v[16 + q], tw[16 * q]);
            dst[(int64_t)(t + 16 * q) * I] = c_add(v[q], w);
            dst[(int64_t)(t + 16 * (q + 16)) * I] = c_sub(v[q], w);
        }
    } else {
        float2* dst = out + o * (int64_t)N * I + i0 + b;
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int n = t + 16 * q;  // stored output index -> cropped row p
            if (n < lo || n >= hi) {
                const float2 w = c_mul(v[16 + (q & 15)], tw[16 * (q & 15)]);
                const float2 x = q < 16 ? c_add(v[q & 15], w) : c_sub(v[q & 15], w);
                const int p = n < lo ? n + N / 2 : n - G + N / 2;
                dst[(int64_t)p * I] = make_float2(x.x, -x.y);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Column passes for G = 16 Q, Q in {5, 6, 8, 9, 10, 12, 15, 18, 20, 24, 25,
// 30, 36, 40} (G = 80-640: the grids of 48-512 point images at oversampling
// 1.25 or 2), the same register four-step as the 256-point
// pass: element n = t + 16k of a column is held by thread t (t = 0..15,
// k = 0..Q-1), which takes the
// Q-point DFT over k and multiplies by exp(-2 pi i t m1 / G); the block
// transposes through LDS (rounds of 16 m1 planes), and thread t takes the
// 16-point DFT over t for m1 = t (+16), giving outputs m1 + Q m2.  The Q-point
// DFTs are Good-Thomas products of two coprime small DFTs (2, 3, 4, 5, 6, 8,
// 9) or Cooley-Tukey 3 x 3 and 5 x 5,
// whose index maps are compile-time register renamings.  Lane = (column b,
// t), b fastest, B consecutive columns per block.
// ---------------------------------------------------------------------------
constexpr int inv_mod(int x, int m) {
    for (int y = 1; y < m; ++y)
        if ((x * y) % m == 1) return y;
    return 1;  // m == 1
}

template <int R>
__device__ __forceinline__ void dft_small(float2* v);

// forward Q-point DFT (Q = A R, gcd(A, R) = 1), in place: input index
// n = (R n1 + A n2) mod Q, output index k = (R (R^-1 mod A) k1 + A (A^-1 mod R) k2) mod Q
template <int A, int R>
__device__ __forceinline__ void dft_pfa(float2* v) {
    constexpr int Q = A * R;
    constexpr int ra = R * inv_mod(R % A, A), ar = A * inv_mod(A % R, R);
    float2 u[A][R];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1)
#pragma unroll
        for (int n2 = 0; n2 < R; ++n2) u[n1][n2] = v[(R * n1 + A * n2) % Q];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) dft_small<R>(u[n1]);
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) {
        float2 c[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) c[n1] = u[n1][k2];
        dft_small<A>(c);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) u[k1][k2] = c[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) v[(ra * k1 + ar * k2) % Q] = u[k1][k2];
}

// cos and sin of 2 pi j / 9 (j = 0..4) and 2 pi j / 25 (j = 0..16): the
// twiddles W_Q^(n2 k1) of the Cooley-Tukey 3 x 3 and 5 x 5 DFTs
__constant__ float2 c_w9[5] = {
    {1.0f, 0.0f}, {0.76604444311897801f, 0.64278760968653925f}, {0.17364817766693041f, 0.98480775301220802f},
    {-0.49999999999999978f, 0.86602540378443871f}, {-0.93969262078590832f, 0.34202014332566888f}};
__constant__ float2 c_w25[17] = {
    {1.0f, 0.0f}, {0.96858316112863108f, 0.24868988716485479f}, {0.87630668004386358f, 0.48175367410171532f},
    {0.72896862742141155f, 0.68454710592868862f}, {0.53582679497899655f, 0.84432792550201508f}, {0.30901699437494745f, 0.95105651629515353f},
    {0.062790519529313527f, 0.99802672842827156f}, {-0.1873813145857246f, 0.98228725072868872f}, {-0.42577929156507272f, 0.90482705246601947f},
    {-0.63742398974868975f, 0.77051324277578925f}, {-0.80901699437494734f, 0.58778525229247325f}, {-0.92977648588825135f, 0.36812455268467814f},
    {-0.99211470131447776f, 0.12533323356430454f}, {-0.99211470131447788f, -0.12533323356430429f}, {-0.92977648588825146f, -0.36812455268467792f},
    {-0.80901699437494778f, -0.58778525229247269f}, {-0.63742398974868952f, -0.77051324277578936f}};

// forward Q-point DFT (Q = A R, A = R = 3 or 5) by Cooley-Tukey: n = R n1 + n2,
// k = k1 + A k2; A-point DFTs over n1, twiddle W_Q^(n2 k1), R-point DFTs over n2
template <int A, int R>
__device__ __forceinline__ void dft_ct(float2* v) {
    static_assert((A == 3 && R == 3) || (A == 5 && R == 5), "twiddle tables for 9 and 25 only");
    const float2* w = A == 3 ? c_w9 : c_w25;
    float2 u[R][A];
#pragma unroll
    for (int n2 = 0; n2 < R; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) u[n2][n1] = v[R * n1 + n2];
        dftR<A>(u[n2]);
#pragma unroll
        for (int k1 = 1; k1 < A; ++k1)
            if (n2 > 0) u[n2][k1] = c_mul(u[n2][k1], make_float2(w[n2 * k1].x, -w[n2 * k1].y));
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        float2 c[R];
#pragma unroll
        for (int n2 = 0; n2 < R; ++n2) c[n2] = u[n2][k1];
        dftR<R>(c);
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) v[k1 + A * k2] = c[k2];
    }
}

// the factors of the Good-Thomas products: 2, 3, 4, 5, 7, 8, 6 = 2 x 3 and 9 = 3 x 3
template <int R>
__device__ __forceinline__ void dft_small(float2* v) {
    if constexpr (R == 6) dft_pfa<2, 3>(v);
    else if constexpr (R == 9) dft_ct<3, 3>(v);
    else dftR<R>(v);
}

// cos and sin of 2 pi m / 32, m = 0..15 (W32^m = c - i s)
__constant__ float2 c_w32[16] = {
    {1.0f, 0.0f},
    {0.98078528040323043f, 0.19509032201612825f},
    {0.92387953251128674f, 0.38268343236508978f},
    {0.83146961230254524f, 0.55557023301960218f},
    {0.70710678118654757f, 0.70710678118654746f},
    {0.55557023301960229f, 0.83146961230254524f},
    {0.38268343236508984f, 0.92387953251128674f},
    {0.19509032201612833f, 0.98078528040323043f},
    {0.0f, 1.0f},
    {-0.19509032201612819f, 0.98078528040323043f},
    {-0.38268343236508973f, 0.92387953251128674f},
    {-0.55557023301960196f, 0.83146961230254546f},
    {-0.70710678118654746f, 0.70710678118654757f},
    {-0.83146961230254535f, 0.55557023301960218f},
    {-0.92387953251128674f, 0.38268343236508989f},
    {-0.98078528040323043f, 0.19509032201612861f}};

// forward 32-point DFT in place: two 16-point DFTs over the even and odd
// elements and a radix-2 step
__device__ __forceinline__ void dft32(float2* v) {
    float2 e[16], o[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        e[s] = v[2 * s];
        o[s] = v[2 * s + 1];
    }
    dft16(e);
    dft16(o);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const float2 w = c_mul(o[m], make_float2(c_w32[m].x, -c_w32[m].y));
        v[m] = c_add(e[m], w);
        v[m + 16] = c_sub(e[m], w);
    }
}

template <int Q>
__device__ __forceinline__ void dftQ(float2* v) {
    if constexpr (Q == 32) dft32(v);
    else if constexpr (Q == 5) dft5(v);
    else if constexpr (Q == 6) dft_pfa<2, 3>(v);
    else if constexpr (Q == 8) dft8(v);
    else if constexpr (Q == 10) dft_pfa<2, 5>(v);
    else if constexpr (Q == 12) dft_pfa<4, 3>(v);
    else if constexpr (Q == 15) dft_pfa<3, 5>(v);
    else if constexpr (Q == 20) dft_pfa<4, 5>(v);
    else if constexpr (Q == 24) dft_pfa<8, 3>(v);
    else if constexpr (Q == 30) dft_pfa<6, 5>(v);
    else if constexpr (Q == 40) dft_pfa<8, 5>(v);
    else if constexpr (Q == 9) dft_ct<3, 3>(v);
    else if constexpr (Q == 18) dft_pfa<2, 9>(v);
    else if constexpr (Q == 25) dft_ct<5, 5>(v);
    else if constexpr (Q == 36) dft_pfa<4, 9>(v);
    else if constexpr (Q == 45) dft_pfa<9, 5>(v);
    else if constexpr (Q == 7) dft7(v);
    else if constexpr (Q == 14) dft_pfa<2, 7>(v);
    else if constexpr (Q == 21) dft_pfa<3, 7>(v);
    else if constexpr (Q == 28) dft_pfa<4, 7>(v);
    else dft_pfa<5, 7>(v);  // 35
}

// the second stage's DFT over the T threads of a line
template <int T>
__device__ __forceinline__ void dftT(float2* v) {
    if constexpr (T == 16) dft16(v);
    else dft8(v);
}

constexpr int CQ_B = 32;  // columns per block (512 threads)

// G = 8 Q (Q = 5, 7, 9, 15, 21, 25, 35, 45) with register passes of 8 threads per line
static inline bool colq8_supported(int G) {
    return G == 40 || G == 56 || G == 72 || G == 120 || G == 168 || G == 200 || G == 280 || G == 360;
}

// G with a register column pass (16 Q, Q as above and 30, 40; 256 and 512 have their own)
static inline bool colq_supported(int G) {
    return G == 80 || G == 96 || G == 112 || G == 128 || G == 144 || G == 160 || G == 192 || G == 224 ||
           G == 240 || G == 288 || G == 320 || G == 336 || G == 384 || G == 400 || G == 448 || G == 480 ||
           G == 576 || G == 640;
}

template <bool FWD, int Q, int B, int T = 16>
__global__ void __launch_bounds__(T * B, Q <= 24 ? 4 : 2) k_fftcolq(int N, int64_t I, const float2* __restrict__ twg,
                                                  const float2* __restrict__ in, float2* __restrict__ out,
                                                  ColMask cm) {
    constexpr int G = T * Q;
    constexpr int NP = Q < T ? Q : T;  // m1 planes per transpose round
    constexpr int PLANE = T * B;         // a 32-lane access covers 32 consecutive columns: no bank conflicts
    __shared__ float2 sh[NP * PLANE];
    __shared__ float2 tw[G];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i0 = (blk % nchunk) * B;
    const bool bok = i0 + b < I;
    const bool zero = !bok || cm_dead_col(cm, i0 + b);  // no column, or one known to be zero
    for (int k = threadIdx.x; k < G; k += T * B) tw[k] = twg[k];
    const int lo = N - N / 2, hi = G - N / 2;  // stored rows [lo, hi) carry no image data
    float2 v[Q];
    if (FWD) {
        const float2* src = in + o * (int64_t)N * I + i0 + b;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const int n = t + T * k;  // stored row
            const int p = n < lo ? n + N / 2 : n - G + N / 2;
            v[k] = (!zero && (n < lo || n >= hi)) ? src[(int64_t)p * I] : make_float2(0.f, 0.f);
        }
    } else {
        const float2* src = in + o * (int64_t)G * I + i0 + b;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const float2 x = !zero ? src[(int64_t)(t + T * k) * I] : make_float2(0.f, 0.f);
            v[k] = make_float2(x.x, -x.y);
        }
    }
    dftQ<Q>(v);
    __syncthreads();  // twiddle table
#pragma unroll
    for (int m = 1; m < Q; ++m) v[m] = c_mul(v[m], tw[t * m]);
#pragma unroll
    for (int h = 0; h < (Q + T - 1) / T; ++h) {
        if (h) __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int m = T * h; m < Q && m < T * h + T; ++m) sh[(m - T * h) * PLANE + t * B + b] = v[m];
        __syncthreads();
        const int m1 = T * h + t;
        if (m1 < Q) {
            float2 u[T];  // u[t'] = twiddled Q-point output m1 of thread t'
#pragma unroll
            for (int tp = 0; tp < T; ++tp) u[tp] = sh[t * PLANE + tp * B + b];
            dftT<T>(u);  // u[m2] = X[m1 + Q m2]
            if (bok) {
                if (FWD) {
                    float2* dst = out + o * (int64_t)G * I + i0 + b;
#pragma unroll
                    for (int m2 = 0; m2 < T; ++m2) dst[(int64_t)(m1 + Q * m2) * I] = u[m2];
                } else {
                    float2* dst = out + o * (int64_t)N * I + i0 + b;
#pragma unroll
                    for (int m2 = 0; m2 < T; ++m2) {
                        const int n = m1 + Q * m2;  // stored output index -> cropped row p
                        if (n < lo || n >= hi) {
                            const int p = n < lo ? n + N / 2 : n - G + N / 2;
                            dst[(int64_t)p * I] = make_float2(u[m2].x, -u[m2].y);
                        }
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// X pass for G = 256 and one B0 segment (any subspace K and batch): the same
// register four-step FFT, 16 columns = (image row, k*Cv + cv) per block.  The
// image side is x-contiguous, the grid side has the columns innermost, so the
// lane map differs between the stages: forward loads with the 16 threads of a
// column on consecutive x and stores with consecutive columns; the adjoint
// the other way round.  The transpose buffer uses a 17-entry row stride so
// both maps are free of LDS bank conflicts.
// ---------------------------------------------------------------------------
constexpr int X256_ROW = 17;  // 16-column transpose strides (Cartesian row passes)
constexpr int X256_PLANE = 16 * X256_ROW + 4;
// columns per block of the row passes, forward and inverse (32 measured no
// faster than 16 on the 3-D headline: 2.59 / 2.31 ms vs 2.57 / 2.29)
constexpr int X256_BF = 16;
constexpr int X256_BA = 16;

// BX columns per block (16 or 32): 16 x BX threads.  M1F (forward): store
// with the 16 threads of a column on consecutive lanes, as the inverse loads:
// where a row has few columns (KC < 8, e.g. the coil-pair grid's 2 values
// per point) the columns of one row then fill whole contiguous runs, while
// columns on consecutive lanes write KC x 8-byte pieces of separate rows.
template <bool FWD, int BX, bool M1F = false>
__global__ void __launch_bounds__(16 * BX, PF_FFT_WPE * 16 / BX) k_fftx256(Geo g, Batch bt,
                                                                         const float2* __restrict__ twg,
                                                                         const float* __restrict__ apod,
                                                                         const float2* __restrict__ in,
                                                                         float2* __restrict__ out, int KC, int YG,
                                                                         int nkc, float factor) {
    constexpr int G = 256;
    constexpr int ROW = BX + 1, PLANE = 16 * ROW + 4;  // transpose strides (float2)
    __shared__ float2 sh[16 * PLANE];
    __shared__ float2 tw[G];
    const int N = g.N[2];
    const int RC = bt.R * bt.Cv;
    const int KCv = bt.K * bt.Cv;
    const int64_t nrows = (int64_t)g.N[0] * g.N[1];
    const int64_t rows_tot = (int64_t)bt.nfg * bt.Cb * nrows;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t row0 = (blk / nkc) * YG;
    const int kc0 = (int)(blk % nkc) * KC;
    const float* apz = apod;
    const float* apy = apod + g.N[0];
    const float* apx = apod + g.N[0] + g.N[1];
    if (threadIdx.x < G) tw[threadIdx.x] = twg[threadIdx.x];
    const int lo = N - N / 2, hi = G - N / 2;
    // column b -> image row and (k, cv) pair
    auto column = [&](int b, int64_t& row, int& kc) {
        const int yy = b / KC;
        kc = kc0 + b % KC;
        row = row0 + yy;
        return yy < YG && row < rows_tot && kc < KCv;
    };
    // image plane of (row, kc) and the row's deapodisation
    auto image_row = [&](int64_t row, int kc, float& ayz) -> int64_t {
        int fg, cb;
        int64_t zy;
        x_row_decode(bt, row, nrows, fg, cb, zy);
        const int k = kc / bt.Cv, cv = kc % bt.Cv;
        const int z = (int)(zy / g.N[1]), y = (int)(zy % g.N[1]);
        ayz = apz[z] * apy[y];
        const int ii = bt.nfg > 1 ? fg : (bt.nimg > 1 ? k : 0);
        const int c = cb * bt.Cv + cv;
        return ((int64_t)c * bt.nimg + ii) * g.ntot + zy * N;
    };
    float2 v[16];
    {
        const int t = FWD ? (threadIdx.x & 15) : (threadIdx.x / BX);
        const int b = FWD ? (threadIdx.x >> 4) : (threadIdx.x % BX);
        int64_t row;
        int kc;
        const bool ok = column(b, row, kc);
        if (FWD) {
            float ayz = 0.f;
            const int64_t base = ok ? image_row(row, kc, ayz) : 0;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int n = t + 16 * k;
                const int p = n < lo ? n + N / 2 : n - G + N / 2;
                v[k] = (ok && (n < lo || n >= hi)) ? c_scale(in[base + p], ayz * apx[p]) : make_float2(0.f, 0.f);
            }
        } else {
            const float2* src = in + (ok ? row : 0) * G * RC + kc;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const float2 x = ok ? src[(int64_t)(t + 16 * k) * RC] : make_float2(0.f, 0.f);
                v[k] = make_float2(x.x, -x.y);
            }
        }
        dft16(v);
        __syncthreads();  // twiddle table
#pragma unroll
        for (int m1 = 1; m1 < 16; ++m1) v[m1] = c_mul(v[m1], tw[t * m1]);
#pragma unroll
        for (int m1 = 0; m1 < 16; ++m1) sh[m1 * PLANE + t * ROW + b] = v[m1];
    }
    __syncthreads();
    const int m1 = (FWD && !M1F) ? (threadIdx.x / BX) : (threadIdx.x & 15);
    const int b = (FWD && !M1F) ? (threadIdx.x % BX) : (threadIdx.x >> 4);
#pragma unroll
    for (int tp = 0; tp < 16; ++tp) v[tp] = sh[m1 * PLANE + tp * ROW + b];
    dft16(v);  // v[m2] = X[m1 + 16 m2] of column b
    int64_t row;
    int kc;
    if (!column(b, row, kc)) return;
    if (FWD) {
        float2* dst = out + row * G * RC + kc;
#pragma unroll
        for (int m2 = 0; m2 < 16; ++m2) dst[(int64_t)(m1 + 16 * m2) * RC] = v[m2];
    } else {
        float ayz = 0.f;
        float2* dst = out + image_row(row, kc, ayz);
        const float sc = factor * ayz;
#pragma unroll
        for (int m2 = 0; m2 < 16; ++m2) {
            const int n = m1 + 16 * m2;
            if (n < lo || n >= hi) {
                const int p = n < lo ? n + N / 2 : n - G + N / 2;
                dst[p] = c_scale(make_float2(v[m2].x, -v[m2].y), sc * apx[p]);
            }
        }
    }
}

// X pass for G = 16 Q (the register column passes' sizes, and the inverse
// at Q = 32): k_fftx256's lane maps and block shape with the 16 x Q
// four-step of k_fftcolq (Q values per thread, the transpose in rounds of 16
// m1 planes).  B0 segments: the forward's columns are (row, v) with
// v = l*K*Cv + kc, each loading its image row times segment l's map; the
// inverse (SEG) keeps all L segments of KC (k, cv) pairs of YG rows in a
// block, writes each segment's cropped row to LDS and sums the segments
// (times the conjugate maps) before one x-contiguous image write.
template <bool FWD, int Q, int BX, bool M1F = false, bool SEG = false, int T = 16>
__global__ void __launch_bounds__(T * BX) k_fftxq(Geo g, Batch bt, const float2* __restrict__ twg,
                                                 const float* __restrict__ apod, const float2* __restrict__ in,
                                                 float2* __restrict__ out, int KC, int YG, int nkc, float factor) {
    constexpr int G = T * Q;
    constexpr int NP = Q < T ? Q : T;
    // transpose strides (float2): as k_fftx256 at T = 16; at T = 8 rows of BX + 4 so the 8 x 4
    // (t, column) lanes of a half-wave hit distinct banks
    constexpr int ROW = T == 16 ? BX + 1 : BX + 4, PLANE = T * ROW + 4;
    __shared__ float2 sh[NP * PLANE];
    __shared__ float2 tw[G];
    __shared__ float2 S[SEG ? BX * G : 1];  // SEG: cropped segment rows, S[(yy, kcl)][l][p]
    const int N = g.N[2];
    const int RC = bt.R * bt.Cv;
    const int KCv = bt.K * bt.Cv;
    const int L = SEG ? bt.L : 1;
    const int VT = FWD ? bt.L * KCv : KCv;  // forward: columns v = l*KCv + kc
    const int64_t nrows = (int64_t)g.N[0] * g.N[1];
    const int64_t rows_tot = (int64_t)bt.nfg * bt.Cb * nrows;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t row0 = (blk / nkc) * YG;
    const int kc0 = (int)(blk % nkc) * KC;
    const float* apz = apod;
    const float* apy = apod + g.N[0];
    const float* apx = apod + g.N[0] + g.N[1];
    for (int k = threadIdx.x; k < G; k += T * BX) tw[k] = twg[k];
    const int lo = N - N / 2, hi = G - N / 2;
    // column b -> row and v (forward: v = l*KCv + kc; inverse: the pair kc,
    // SEG: segment l of pair kc)
    auto column = [&](int b, int64_t& row, int& v) {
        const int lkc = L * KC;
        const int yy = b / lkc, r = b % lkc;
        const int kc = kc0 + r % KC;
        v = (r / KC) * KCv + kc;
        row = row0 + yy;
        return yy < YG && row < rows_tot && kc < VT;
    };
    auto image_row = [&](int64_t row, int kc, float& ayz, int64_t& zy) -> int64_t {
        int fg, cb;
        x_row_decode(bt, row, nrows, fg, cb, zy);
        const int k = kc / bt.Cv, cv = kc % bt.Cv;
        const int z = (int)(zy / g.N[1]), y = (int)(zy % g.N[1]);
        ayz = apz[z] * apy[y];
        const int ii = bt.nfg > 1 ? fg : (bt.nimg > 1 ? k : 0);
        const int c = cb * bt.Cv + cv;
        return ((int64_t)c * bt.nimg + ii) * g.ntot + zy * N;
    };
    float2 v[Q];
    const int t = FWD ? (threadIdx.x % T) : (threadIdx.x / BX);
    {
        const int b = FWD ? (threadIdx.x / T) : (threadIdx.x % BX);
        int64_t row;
        int vc;
        const bool ok = column(b, row, vc);
        if (FWD) {
            float ayz = 0.f;
            int64_t zy = 0;
            const int kc = vc % KCv, l = vc / KCv;
            const int64_t base = ok ? image_row(row, kc, ayz, zy) : 0;
            const float2* b0 = (ok && bt.b0s) ? bt.b0s + (int64_t)l * g.ntot + zy * N : nullptr;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int n = t + T * k;
                const int p = n < lo ? n + N / 2 : n - G + N / 2;
                float2 x = make_float2(0.f, 0.f);
                if (ok && (n < lo || n >= hi)) {
                    x = c_scale(in[base + p], ayz * apx[p]);
                    if (b0) x = c_mul(x, b0[p]);
                }
                v[k] = x;
            }
        } else {
            const float2* src = in + (ok ? row : 0) * G * RC + vc;
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const float2 x = ok ? src[(int64_t)(t + T * k) * RC] : make_float2(0.f, 0.f);
                v[k] = make_float2(x.x, -x.y);
            }
        }
        dftQ<Q>(v);
        __syncthreads();  // twiddle table
#pragma unroll
        for (int m = 1; m < Q; ++m) v[m] = c_mul(v[m], tw[t * m]);
    }
    const int sb = FWD ? (threadIdx.x / T) : (threadIdx.x % BX);  // stage-1 column of this thread
    const int j = (FWD && !M1F) ? (threadIdx.x / BX) : (threadIdx.x % T);
    const int b = (FWD && !M1F) ? (threadIdx.x % BX) : (threadIdx.x / T);
    int64_t row;
    int vc;
    const bool ok = column(b, row, vc);
#pragma unroll
    for (int h = 0; h < (Q + T - 1) / T; ++h) {
        if (h) __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int m = T * h; m < Q && m < T * h + T; ++m) sh[(m - T * h) * PLANE + t * ROW + sb] = v[m];
        __syncthreads();
        const int m1 = T * h + j;
        if (m1 < Q && ok) {
            float2 u[T];
#pragma unroll
            for (int tp = 0; tp < T; ++tp) u[tp] = sh[j * PLANE + tp * ROW + b];
            dftT<T>(u);  // u[m2] = X[m1 + Q m2] of column b
            if (FWD) {
                float2* dst = out + row * G * RC + vc;
#pragma unroll
                for (int m2 = 0; m2 < T; ++m2) dst[(int64_t)(m1 + Q * m2) * RC] = u[m2];
            } else if (SEG) {
                const int lkc = L * KC, r = b % lkc;  // b = (yy, l, kcl) -> segment row (yy*KC + kcl)*L + l
                float2* Sr = S + ((int64_t)((b / lkc) * KC + r % KC) * L + r / KC) * N;
#pragma unroll
                for (int m2 = 0; m2 < T; ++m2) {
                    const int n = m1 + Q * m2;
                    if (n < lo || n >= hi) Sr[n < lo ? n + N / 2 : n - G + N / 2] = make_float2(u[m2].x, -u[m2].y);
                }
            } else {
                float ayz = 0.f;
                int64_t zy;
                float2* dst = out + image_row(row, vc, ayz, zy);
                const float sc = factor * ayz;
#pragma unroll
                for (int m2 = 0; m2 < T; ++m2) {
                    const int n = m1 + Q * m2;
                    if (n < lo || n >= hi) {
                        const int p = n < lo ? n + N / 2 : n - G + N / 2;
                        dst[p] = c_scale(make_float2(u[m2].x, -u[m2].y), sc * apx[p]);
                    }
                }
            }
        }
    }
    if constexpr (SEG) {
        __syncthreads();
        const int NC = YG * KC;  // (row, kc) image rows of this block
        for (int it = threadIdx.x; it < NC * N; it += T * BX) {
            const int p = it % N, cc = it / N;
            const int64_t rw = row0 + cc / KC;
            const int kc = kc0 + cc % KC;
            if (rw >= rows_tot || kc >= KCv) continue;
            float ayz = 0.f;
            int64_t zy;
            const int64_t base = image_row(rw, kc, ayz, zy);
            const float2* Sp = S + (int64_t)cc * L * N + p;
            float2 acc = make_float2(0.f, 0.f);
            if (bt.b0s) {
                const float2* b0 = bt.b0s + zy * N + p;
                for (int l = 0; l < L; ++l) acc = c_add(acc, c_mulc(Sp[l * N], b0[(int64_t)l * g.ntot]));
            } else {
                for (int l = 0; l < L; ++l) acc = c_add(acc, Sp[l * N]);
            }
            out[base + p] = c_scale(acc, factor * ayz * apx[p]);
        }
    }
}

// ---------------------------------------------------------------------------
// X pass for G = 512 with any number of B0 segments L (<= X512_B) and any
// subspace K: a register four-step FFT with 512 = 32 x 16 (32 threads per
// column hold 16 values each for a 16-point DFT, twiddle W512^(t m1), one LDS
// transpose, then threads m1 < 16 take the 32-point DFT over t as two
// 16-point DFTs and a radix-2 step), X512_B columns per block.
//
// Forward: block columns are X512_B consecutive grid columns (row, v) with
// v = l*K*Cv + kc, so every output row n is one contiguous run; a column
// loads its image row (deapodised, times the segment's B0 map) with the 32
// threads on consecutive x.
// Adjoint: a block takes all L segments of KC (k, cv) pairs (YG rows when
// L*K*Cv is smaller than the block), loads with the columns on consecutive
// lanes, and after the FFT sums the segments (times the conjugate maps)
// through LDS, so the image rows are written once, x on consecutive lanes.
// M1F (forward, rows of at most 2 values per point, e.g. the coil-pair
// grid): the second stage puts a column's 16 threads m1 on consecutive lanes,
// so the 2 columns of a row write 256-byte runs instead of the 64-byte pieces
// of 8 rows that columns on consecutive lanes give.
// LDS strides: ds_read/write_b64 serve 32 lanes per cycle; the forward's
// transpose writes have t on consecutive lanes (odd row stride) and its reads
// columns (plane stride = 16 mod 32) or, M1F, m1 (plane stride = 2 mod 32);
// the adjoint's writes columns (row stride 16 with two rows per 32 lanes) and
// its reads m1 (plane stride = 2 mod 32).
// ---------------------------------------------------------------------------
constexpr int X512_B = 16;

template <bool FWD, bool M1F = false>
struct X512Strides {
    static constexpr int ROW = FWD ? X512_B + 1 : X512_B;
    static constexpr int PLANE = (FWD && !M1F) ? 32 * ROW + 16 : 32 * ROW + 2;
};

// image plane base (x = 0 of the row) of (row, kc); zy = the row's z*Ny + y
__device__ __forceinline__ int64_t x_image_row(const Geo& g, const Batch& bt, int64_t row, int64_t nrows, int kc,
                                               const float* apod, float& ayz, int64_t& zy) {
    int fg, cb;
    x_row_decode(bt, row, nrows, fg, cb, zy);
    const int k = kc / bt.Cv, cv = kc % bt.Cv;
    const int z = (int)(zy / g.N[1]), y = (int)(zy % g.N[1]);
    ayz = apod[z] * apod[g.N[0] + y];
    const int ii = bt.nfg > 1 ? fg : (bt.nimg > 1 ? k : 0);
    const int c = cb * bt.Cv + cv;
    return ((int64_t)c * bt.nimg + ii) * g.ntot + zy * g.N[2];
}

template <bool M1F>
__global__ void __launch_bounds__(32 * X512_B) k_fftx512_fwd(Geo g, Batch bt, const float2* __restrict__ twg,
                                                            const float* __restrict__ apod,
                                                            const float2* __restrict__ in,
                                                            float2* __restrict__ out) {
    constexpr int G = 512, BX = X512_B;
    constexpr int ROW = X512Strides<true, M1F>::ROW, PLANE = X512Strides<true, M1F>::PLANE;
    __shared__ float2 sh[16 * PLANE];
    __shared__ float2 tw[G];
    const int N = g.N[2];
    const int RC = bt.R * bt.Cv;
    const int KCv = bt.K * bt.Cv;
    const int64_t nrows = (int64_t)g.N[0] * g.N[1];
    const int64_t ncols = (int64_t)bt.nfg * bt.Cb * nrows * RC;
    const int64_t c0 = pf_xcd_remap(blockIdx.x, gridDim.x) * BX;
    const float* apx = apod + g.N[0] + g.N[1];
    for (int k = threadIdx.x; k < G; k += 32 * BX) tw[k] = twg[k];
    const int lo = N - N / 2, hi = G - N / 2;  // stored positions [lo, hi) hold no image data
    float2 v[32];
    {
        const int t = threadIdx.x & 31, b = threadIdx.x >> 5;
        const int64_t gcol = c0 + b;
        const bool ok = gcol < ncols;
        const int64_t row = ok ? gcol / RC : 0;
        const int vr = ok ? (int)(gcol % RC) : 0;
        const int l = vr / KCv, kc = vr % KCv;
        float ayz = 0.f;
        int64_t zy = 0;
        const int64_t base = x_image_row(g, bt, row, nrows, kc, apod, ayz, zy);
        const float2* b0 = bt.b0s ? bt.b0s + (int64_t)l * g.ntot + zy * N : nullptr;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int n = t + 32 * k;
            const int p = n < lo ? n + N / 2 : n - G + N / 2;
            float2 x = make_float2(0.f, 0.f);
            if (ok && (n < lo || n >= hi)) {
                x = c_scale(in[base + p], ayz * apx[p]);
                if (b0) x = c_mul(x, b0[p]);
            }
            v[k] = x;
        }
        dft16(v);
        __syncthreads();  // twiddle table
#pragma unroll
        for (int m1 = 1; m1 < 16; ++m1) v[m1] = c_mul(v[m1], tw[t * m1]);
#pragma unroll
        for (int m1 = 0; m1 < 16; ++m1) sh[m1 * PLANE + t * ROW + b] = v[m1];
    }
    __syncthreads();
    if (threadIdx.x >= 16 * BX) return;
    const int m1 = M1F ? (threadIdx.x & 15) : (threadIdx.x / BX);
    const int b = M1F ? (threadIdx.x >> 4) : (threadIdx.x % BX);
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        v[s] = sh[m1 * PLANE + (2 * s) * ROW + b];
        v[16 + s] = sh[m1 * PLANE + (2 * s + 1) * ROW + b];
    }
    dft16(v);
    dft16(v + 16);
    const int64_t gcol = c0 + b;
    if (gcol >= ncols) return;
    float2* dst = out + (gcol / RC) * G * RC + gcol % RC;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const float2 w = c_mul(v[16 + q], tw[16 * q]);
        dst[(int64_t)(m1 + 16 * q) * RC] = c_add(v[q], w);
        dst[(int64_t)(m1 + 16 * (q + 16)) * RC] = c_sub(v[q], w);
    }
}

__global__ void __launch_bounds__(32 * X512_B) k_fftx512_adj(Geo g, Batch bt, const float2* __restrict__ twg,
                                                            const float* __restrict__ apod,
                                                            const float2* __restrict__ in,
                                                            float2* __restrict__ img, XBlk xb, float factor) {
    constexpr int G = 512, BX = X512_B;
    constexpr int ROW = X512Strides<false>::ROW, PLANE = X512Strides<false>::PLANE;
    static_assert(16 * PLANE >= BX * G, "segment rows must fit the transpose buffer");
    __shared__ float2 sh[16 * PLANE];
    __shared__ float2 tw[G];
    const int N = g.N[2];
    const int RC = bt.R * bt.Cv;
    const int KCv = bt.K * bt.Cv;
    const int L = bt.L;
    const int VC = L * xb.KC;                  // columns per image row
    const int NC = xb.YG * xb.KC;              // (row, kc) image rows per block
    // segment row stride: padded when it fits, so two columns of a half-wave use different banks
    const int NS = BX * (N + 16) <= 16 * PLANE ? N + 16 : N;
    const int64_t nrows = (int64_t)g.N[0] * g.N[1];
    const int64_t rows_tot = (int64_t)bt.nfg * bt.Cb * nrows;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t row0 = (blk / xb.nkc) * xb.YG;
    const int kc0 = (int)(blk % xb.nkc) * xb.KC;
    const float* apx = apod + g.N[0] + g.N[1];
    for (int k = threadIdx.x; k < G; k += 32 * BX) tw[k] = twg[k];
    const int lo = N - N / 2, hi = G - N / 2;
    // column b -> (image row, segment, kc)
    auto column = [&](int b, int64_t& row, int& l, int& kc) {
        const int yy = b / VC, j = b % VC;
        l = j / xb.KC;
        kc = kc0 + j % xb.KC;
        row = row0 + yy;
        return yy < xb.YG && row < rows_tot && kc < KCv;
    };
    float2 v[32];
    {
        const int b = threadIdx.x % BX, t = threadIdx.x / BX;
        int64_t row;
        int l, kc;
        const bool ok = column(b, row, l, kc);
        const float2* src = in + (ok ? row : 0) * G * RC + l * KCv + kc;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const float2 x = ok ? src[(int64_t)(t + 32 * k) * RC] : make_float2(0.f, 0.f);
            v[k] = make_float2(x.x, -x.y);  // inverse via the forward FFT of the conjugate
        }
        dft16(v);
        __syncthreads();  // twiddle table
#pragma unroll
        for (int m1 = 1; m1 < 16; ++m1) v[m1] = c_mul(v[m1], tw[t * m1]);
#pragma unroll
        for (int m1 = 0; m1 < 16; ++m1) sh[m1 * PLANE + t * ROW + b] = v[m1];
    }
    __syncthreads();
    const bool act = threadIdx.x < 16 * BX;
    const int m1 = threadIdx.x & 15, b = (threadIdx.x >> 4) % BX;
    if (act) {
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            v[s] = sh[m1 * PLANE + (2 * s) * ROW + b];
            v[16 + s] = sh[m1 * PLANE + (2 * s + 1) * ROW + b];
        }
        dft16(v);
        dft16(v + 16);
    }
    __syncthreads();  // the transpose buffer becomes the segment buffer S[(yy, kcl)][l][p]
    if (act) {
        int64_t row;
        int l, kc;
        if (column(b, row, l, kc)) {
            float2* S = sh + ((int64_t)((b / VC) * xb.KC + (kc - kc0)) * L + l) * NS;
#pragma unroll
            for (int q = 0; q < 32; ++q) {
                const int n = m1 + 16 * q;
                if (n < lo || n >= hi) {
                    const float2 w = c_mul(v[16 + (q & 15)], tw[16 * (q & 15)]);
                    const float2 x = q < 16 ? c_add(v[q & 15], w) : c_sub(v[q & 15], w);
                    S[n < lo ? n + N / 2 : n - G + N / 2] = make_float2(x.x, -x.y);
                }
            }
        }
    }
    __syncthreads();
    for (int it = threadIdx.x; it < NC * N; it += 32 * BX) {
        const int p = it % N, cc = it / N;
        const int yy = cc / xb.KC, kc = kc0 + cc % xb.KC;
        const int64_t row = row0 + yy;
        if (row >= rows_tot || kc >= KCv) continue;
        float ayz = 0.f;
        int64_t zy = 0;
        const int64_t base = x_image_row(g, bt, row, nrows, kc, apod, ayz, zy);
        const float2* S = sh + (int64_t)cc * L * NS + p;
        float2 acc = make_float2(0.f, 0.f);
        if (bt.b0s) {
            const float2* b0 = bt.b0s + zy * N + p;
            for (int l = 0; l < L; ++l) acc = c_add(acc, c_mulc(S[l * NS], b0[(int64_t)l * g.ntot]));
        } else {
            for (int l = 0; l < L; ++l) acc = c_add(acc, S[l * NS]);
        }
        img[base + p] = c_scale(acc, factor * ayz * apx[p]);
    }
}


// ---------------------------------------------------------------------------
// Cartesian passes (masked centred FFT without basis or B0): the same FFTs
// directly on the caller's (planes, Nz, Ny, Nx) layout, oversampling 1.
// Both sides are centred: stored position n of a line holds centred index
// (n + N/2) mod N, in the image and in k-space alike, so a pass reads and
// writes line element (n + N/2) mod N.  The last forward pass multiplies by
// the mask (frame t = plane mod T) and the norm factor on the way out; the
// first inverse pass applies the mask on the way in and the last one the
// factor.  A block reads all of its lines before it writes, so passes run in
// place.
// ---------------------------------------------------------------------------
struct CartEpi {
    const float* mask;  // (mframes, rows, I) centred, or null
    int T;              // frames per plane group (plane = c * T + t)
    int mframes;
    float factor;
};

__device__ __forceinline__ int crot(int n, int N) {
    const int j = n + N / 2;
    return j >= N ? j - N : j;
}

__device__ __forceinline__ float cmask(const CartEpi& e, int64_t o, int rows, int j, int64_t I, int64_t i) {
    if (!e.mask) return 1.f;
    const int64_t tm = e.mframes > 1 ? o % e.T : 0;
    return e.mask[(tm * rows + j) * I + i];
}

// column pass, G = 256: view (O, 256, I); four-step register FFT as k_fftcol256
template <bool FWD, int B>
__global__ void __launch_bounds__(16 * B, PF_FFT_WPE * 16 / B) k_cartcol256(int64_t I, const float2* __restrict__ twg,
                                                                          const float2* in, float2* out, CartEpi e) {
    constexpr int G = 256;
    constexpr int CART256_PLANE = 16 * B + 8;  // LDS plane stride (float2) of the transpose
    __shared__ float2 sh[16 * CART256_PLANE];
    __shared__ float2 tw[G];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i = (blk % nchunk) * B + b;
    const bool bok = i < I;
    if (threadIdx.x < G) tw[threadIdx.x] = twg[threadIdx.x];
    const float2* src = in + o * G * I + i;
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int j = crot(t + 16 * k, G);
        float2 x = bok ? src[(int64_t)j * I] : make_float2(0.f, 0.f);
        if (!FWD) {
            if (bok) x = c_scale(x, cmask(e, o, G, j, I, i));
            x = make_float2(x.x, -x.y);
        }
        v[k] = x;
    }
    dft16(v);
    __syncthreads();
#pragma unroll
    for (int m1 = 1; m1 < 16; ++m1) v[m1] = c_mul(v[m1], tw[t * m1]);
#pragma unroll
    for (int m1 = 0; m1 < 16; ++m1) sh[m1 * CART256_PLANE + t * B + b] = v[m1];
    __syncthreads();
#pragma unroll
    for (int tp = 0; tp < 16; ++tp) v[tp] = sh[t * CART256_PLANE + tp * B + b];
    dft16(v);  // v[m2] = X[t + 16 m2]
    if (!bok) return;
    float2* dst = out + o * G * I + i;
#pragma unroll
    for (int m2 = 0; m2 < 16; ++m2) {
        const int j = crot(t + 16 * m2, G);
        float2 y = FWD ? v[m2] : make_float2(v[m2].x, -v[m2].y);
        const float sc = FWD ? e.factor * cmask(e, o, G, j, I, i) : e.factor;
        dst[(int64_t)j * I] = c_scale(y, sc);
    }
}

// column pass, G = 16 Q (128-640 points): the four-step of k_fftcolq with
// the centred reads and writes and the mask / factor epilogue of k_cartcol256
template <bool FWD, int Q, int B, int T = 16>
__global__ void __launch_bounds__(T * B) k_cartcolq(int64_t I, const float2* __restrict__ twg, const float2* in,
                                                   float2* out, CartEpi e) {
    constexpr int G = T * Q;
    constexpr int NP = Q < T ? Q : T;
    constexpr int PLANE = T * B;   // a 32-lane access covers 32 consecutive slots: no bank conflicts
    __shared__ float2 sh[NP * PLANE];
    __shared__ float2 tw[G];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i = (blk % nchunk) * B + b;
    const bool bok = i < I;
    for (int k = threadIdx.x; k < G; k += T * B) tw[k] = twg[k];
    const float2* src = in + o * G * I + i;
    float2 v[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int j = crot(t + T * k, G);
        float2 x = bok ? src[(int64_t)j * I] : make_float2(0.f, 0.f);
        if (!FWD) {
            if (bok) x = c_scale(x, cmask(e, o, G, j, I, i));
            x = make_float2(x.x, -x.y);
        }
        v[k] = x;
    }
    dftQ<Q>(v);
    __syncthreads();  // twiddle table
#pragma unroll
    for (int m = 1; m < Q; ++m) v[m] = c_mul(v[m], tw[t * m]);
    float2* dst = out + o * G * I + i;
#pragma unroll
    for (int h = 0; h < (Q + T - 1) / T; ++h) {
        if (h) __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int m = T * h; m < Q && m < T * h + T; ++m) sh[(m - T * h) * PLANE + t * B + b] = v[m];
        __syncthreads();
        const int m1 = T * h + t;
        if (m1 < Q && bok) {
            float2 u[T];
#pragma unroll
            for (int tp = 0; tp < T; ++tp) u[tp] = sh[t * PLANE + tp * B + b];
            dftT<T>(u);  // u[m2] = X[m1 + Q m2]
#pragma unroll
            for (int m2 = 0; m2 < T; ++m2) {
                const int j = crot(m1 + Q * m2, G);
                const float2 y = FWD ? u[m2] : make_float2(u[m2].x, -u[m2].y);
                const float sc = FWD ? e.factor * cmask(e, o, G, j, I, i) : e.factor;
                dst[(int64_t)j * I] = c_scale(y, sc);
            }
        }
    }
}

// column pass, any supported G: Stockham in LDS over B columns
template <bool FWD>
__global__ void __launch_bounds__(FT, PF_FFT_WPE) k_cartcol(AxisFFT A, int64_t I, int B, int bshift, const float2* in,
                                                            float2* out, CartEpi e) {
    extern __shared__ float2 sh[];
    const int G = A.G;
    const int LD = B + 1;
    float2* buf = sh;
    float2* tw = sh + G * LD;
    const int64_t nchunk = (I + B - 1) / B;
    const int64_t blk = pf_xcd_remap(blockIdx.x, gridDim.x);
    const int64_t o = blk / nchunk;
    const int64_t i0 = (blk % nchunk) * B;
    const int bw = (int)std::min<int64_t>(B, I - i0);
    for (int q = threadIdx.x; q < G; q += FT) tw[q] = A.tw[q];
    const float2* src = in + o * (int64_t)G * I + i0;
    for (int q = threadIdx.x; q < (G << bshift); q += FT) {
        const int b = q & (B - 1), n = q >> bshift;
        const int j = crot(n, G);
        float2 x = b < bw ? src[(int64_t)j * I + b] : make_float2(0.f, 0.f);
        if (!FWD) {
            if (b < bw) x = c_scale(x, cmask(e, o, G, j, I, i0 + b));
            x = make_float2(x.x, -x.y);
        }
        buf[n * LD + b] = x;
    }
    __syncthreads();
    lds_fft(buf, LD, B, A, tw);
    float2* dst = out + o * (int64_t)G * I + i0;
    for (int q = threadIdx.x; q < (G << bshift); q += FT) {
        const int b = q & (B - 1), n = q >> bshift;
        if (b >= bw) continue;
        const int j = crot(n, G);
        const float2 x = buf[n * LD + b];
        const float2 y = FWD ? x : make_float2(x.x, -x.y);
        const float sc = FWD ? e.factor * cmask(e, o, G, j, I, i0 + b) : e.factor;
        dst[(int64_t)j * I + b] = c_scale(y, sc);
    }
}

// row pass, G = 256: 16 contiguous rows per block; load lanes run along the
// row, the transpose hands every thread one output phase m1, and the store
// lanes run along the row again
template <bool FWD>
__global__ void __launch_bounds__(256, PF_FFT_WPE) k_cartrow256(int64_t nrows, const float2* __restrict__ twg,
                                                              const float2* in, float2* out, CartEpi e) {
    constexpr int G = 256;
    __shared__ float2 sh[16 * X256_PLANE];
    __shared__ float2 tw[G];
    const int64_t r0 = (int64_t)pf_xcd_remap(blockIdx.x, gridDim.x) * 16;
    tw[threadIdx.x] = twg[threadIdx.x];
    float2 v[16];
    {
        const int t = threadIdx.x & 15, b = threadIdx.x >> 4;
        const int64_t row = r0 + b;
        const bool ok = row < nrows;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int j = crot(t + 16 * k, G);
            float2 x = ok ? in[row * G + j] : make_float2(0.f, 0.f);
            if (!FWD) {
                if (ok) x = c_scale(x, cmask(e, row, 1, 0, G, j));
                x = make_float2(x.x, -x.y);
            }
            v[k] = x;
        }
        dft16(v);
        __syncthreads();
#pragma unroll
        for (int m1 = 1; m1 < 16; ++m1) v[m1] = c_mul(v[m1], tw[t * m1]);
#pragma unroll
        for (int m1 = 0; m1 < 16; ++m1) sh[m1 * X256_PLANE + t * X256_ROW + b] = v[m1];
    }
    __syncthreads();
    const int m1 = threadIdx.x & 15, b = threadIdx.x >> 4;
#pragma unroll
    for (int tp = 0; tp < 16; ++tp) v[tp] = sh[m1 * X256_PLANE + tp * X256_ROW + b];
    dft16(v);  // v[m2] = X[m1 + 16 m2] of row b
    const int64_t row = r0 + b;
    if (row >= nrows) return;
#pragma unroll
    for (int m2 = 0; m2 < 16; ++m2) {
        const int j = crot(m1 + 16 * m2, G);
        const float2 y = FWD ? v[m2] : make_float2(v[m2].x, -v[m2].y);
        const float sc = FWD ? e.factor * cmask(e, row, 1, 0, G, j) : e.factor;
        out[row * G + j] = c_scale(y, sc);
    }
}

// row pass, G = T Q (T = 16 or 8 threads per row): 256 / T rows per block,
// k_cartrow256's lane maps with the T x Q four-step (transpose rounds of T m1
// planes)
template <bool FWD, int Q, int T = 16>
__global__ void __launch_bounds__(256) k_cartrowq(int64_t nrows, const float2* __restrict__ twg, const float2* in,
                                                float2* out, CartEpi e) {
    constexpr int G = T * Q;
    constexpr int NP = Q < T ? Q : T;
    constexpr int RB = 256 / T;                                 // rows per block
    constexpr int ROW = T == 16 ? RB + 1 : RB + 4, PLANE = T * ROW + 4;  // as k_fftxq
    __shared__ float2 sh[NP * PLANE];
    __shared__ float2 tw[G];
    const int64_t r0 = (int64_t)pf_xcd_remap(blockIdx.x, gridDim.x) * RB;
    for (int k = threadIdx.x; k < G; k += 256) tw[k] = twg[k];
    const int t = threadIdx.x % T, b = threadIdx.x / T;
    const int64_t row = r0 + b;
    const bool ok = row < nrows;
    float2 v[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int j = crot(t + T * k, G);
        float2 x = ok ? in[row * G + j] : make_float2(0.f, 0.f);
        if (!FWD) {
            if (ok) x = c_scale(x, cmask(e, row, 1, 0, G, j));
            x = make_float2(x.x, -x.y);
        }
        v[k] = x;
    }
    dftQ<Q>(v);
    __syncthreads();  // twiddle table
#pragma unroll
    for (int m = 1; m < Q; ++m) v[m] = c_mul(v[m], tw[t * m]);
#pragma unroll
    for (int h = 0; h < (Q + T - 1) / T; ++h) {
        if (h) __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int m = T * h; m < Q && m < T * h + T; ++m) sh[(m - T * h) * PLANE + t * ROW + b] = v[m];
        __syncthreads();
        const int m1 = T * h + t;  // second stage: thread t takes m1 = 16 h + t of row b
        if (m1 < Q && ok) {
            float2 u[T];
#pragma unroll
            for (int tp = 0; tp < T; ++tp) u[tp] = sh[t * PLANE + tp * ROW + b];
            dftT<T>(u);  // u[m2] = X[m1 + Q m2] of row b
#pragma unroll
            for (int m2 = 0; m2 < T; ++m2) {
                const int j = crot(m1 + Q * m2, G);
                const float2 y = FWD ? u[m2] : make_float2(u[m2].x, -u[m2].y);
                const float sc = FWD ? e.factor * cmask(e, row, 1, 0, G, j) : e.factor;
                out[row * G + j] = c_scale(y, sc);
            }
        }
    }
}

// row pass, any supported G: B rows per block transposed into LDS columns
template <bool FWD>
__global__ void __launch_bounds__(FT, PF_FFT_WPE) k_cartrow(AxisFFT A, int64_t nrows, int B, const float2* in,
                                                            float2* out, CartEpi e) {
    extern __shared__ float2 sh[];
    const int G = A.G;
    const int LD = B + 1;
    float2* buf = sh;
    float2* tw = sh + G * LD;
    const int64_t r0 = (int64_t)pf_xcd_remap(blockIdx.x, gridDim.x) * B;
    for (int q = threadIdx.x; q < G; q += FT) tw[q] = A.tw[q];
    for (int q = threadIdx.x; q < B * G; q += FT) {
        const int j = q % G, r = q / G;  // centred element j of row r
        const int64_t row = r0 + r;
        float2 x = row < nrows ? in[row * G + j] : make_float2(0.f, 0.f);
        if (!FWD) {
            if (row < nrows) x = c_scale(x, cmask(e, row, 1, 0, G, j));
            x = make_float2(x.x, -x.y);
        }
        buf[stored_pos(j, G, G) * LD + r] = x;
    }
    __syncthreads();
    lds_fft(buf, LD, B, A, tw);
    for (int q = threadIdx.x; q < B * G; q += FT) {
        const int j = q % G, r = q / G;
        const int64_t row = r0 + r;
        if (row >= nrows) continue;
        const float2 x = buf[stored_pos(j, G, G) * LD + r];
        const float2 y = FWD ? x : make_float2(x.x, -x.y);
        const float sc = FWD ? e.factor * cmask(e, row, 1, 0, G, j) : e.factor;
        out[row * G + j] = c_scale(y, sc);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct pf_fftk_axis {
    int G = 0, N = 0;
    std::vector<int> rad;
    float2* tw = nullptr;
};

struct pf_fftk {
    pf_fftk_axis ax[3];
};

static bool factorize(int G, std::vector<int>& rad) {
    rad.clear();
    int g = G;
    for (int r : {8, 5, 4, 3, 2})
        while (g % r == 0) {
            rad.push_back(r);
            g /= r;
        }
    return g == 1 && (int)rad.size() <= 16;
}

// the error of a pass that needs the Stockham kernels on an axis without a radix set
constexpr const char* kNoStockham =
    "this grid size (a factor 7) runs only the register FFT passes, which take at most 16 field-map segments: "
    "use a segment batch size of at most 16";

// an axis the fused passes can transform: a Stockham radix set (8, 5, 4, 3, 2),
// or a register four-step size (the 7-smooth ones have no Stockham kernel)
static bool fftk_axis_ok(int G) {
    std::vector<int> rad;
    return factorize(G, rad) || (PF_FFTQ && (colq_supported(G) || colq8_supported(G)));
}

bool pf_fftk_supported(const pf_plan* p) {
    for (int a = 3 - p->ndim; a < 3; ++a) {
        if (p->os[a] > GBMAX || !fftk_axis_ok(p->os[a])) return false;
        if (p->shape[a] > p->os[a]) return false;
    }
    return true;
}

int pf_fftk_init(pf_plan* p) {
    if (p->fftk) return PF_OK;
    pf_fftk* f = new pf_fftk();
    for (int a = 3 - p->ndim; a < 3; ++a) {
        pf_fftk_axis& X = f->ax[a];
        X.G = p->os[a];
        X.N = p->shape[a];
        if (!factorize(X.G, X.rad)) {
            X.rad.clear();  // register passes only (a factor 7)
            if (!fftk_axis_ok(X.G)) {
                pf_fftk_destroy(f);
                return pf_fail(PF_ERR_UNSUPPORTED, "grid size has prime factors other than 2, 3, 5");
            }
        }
        std::vector<float2> h(X.G);
        for (int k = 0; k < X.G; ++k) {
            double ang = -2.0 * M_PI * (double)k / (double)X.G;
            h[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        if (hipMalloc(&X.tw, X.G * sizeof(float2)) != hipSuccess ||
            hipMemcpy(X.tw, h.data(), X.G * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
            pf_fftk_destroy(f);
            return pf_fail(PF_ERR_HIP, "twiddle table upload failed");
        }
    }
    if ((size_t)(GBMAX + GBMAX / 8 + 4096) * sizeof(float2) > 64 * 1024) {
        // blocks of more than 64 KB of LDS must be allowed explicitly
        const void* kernels[] = {reinterpret_cast<const void*>(k_fftx_fwd), reinterpret_cast<const void*>(k_fftx_adj),
                                 reinterpret_cast<const void*>(k_fftcol<true>),
                                 reinterpret_cast<const void*>(k_fftcol<false>)};
        for (const void* k : kernels)
            if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess) {
                pf_fftk_destroy(f);
                return pf_fail(PF_ERR_HIP, "could not raise the FFT kernels' LDS limit");
            }
    }
    p->fftk = f;
    return PF_OK;
}

// The fused-FFT state (twiddle tables, raised LDS limits) is built once when a
// plan or operator is created; operator calls only check that it is there, so
// they neither allocate nor synchronise and may be captured into a graph.
int pf_fftk_ready(const pf_plan* p) {
    if (p->fftk) return PF_OK;
    return pf_fail(PF_ERR_ARG, "plan has no fused-FFT state (it was not built by a pf_*_create call)");
}

void pf_fftk_destroy(pf_fftk* f) {
    if (!f) return;
    for (auto& X : f->ax)
        if (X.tw) (void)hipFree(X.tw);
    delete f;
}

static AxisFFT axis_desc(const pf_fftk_axis& X) {
    AxisFFT A;
    A.G = X.G;
    A.N = X.N;
    A.nrad = (int)X.rad.size();
    int ns = 1;
    for (int i = 0; i < A.nrad; ++i) {
        A.rad[i] = X.rad[i];
        A.nsm[i] = magic_of(ns);
        ns *= X.rad[i];
    }
    A.tw = X.tw;
    return A;
}

static Geo geo_of(const pf_plan* p) {
    Geo g;
    for (int a = 0; a < 3; ++a) {
        g.N[a] = p->shape[a];
        g.G[a] = p->os[a];
    }
    g.ntot = pf_ntot(p);
    return g;
}

static Batch batch_of(const pf_batch& b) {
    Batch t{b.C, b.Cv, b.Cb, b.R, b.L, b.K, b.nfg, b.nimg, b.b0s};
    return t;
}

// columns per block for a line of length G
static int cols_for(int G) {
    int B = 1;
    while (B * 2 * G <= GBMAX && B < 64) B *= 2;
    return B;
}

size_t pf_fftk_extra_bytes(const pf_plan* p, const pf_batch& b) {
    const size_t RC = (size_t)b.R * b.Cv;
    const size_t blocks = (size_t)b.nfg * b.Cb * RC * sizeof(float2);
    if (p->ndim == 3) return blocks * p->shape[0] * p->os[1] * p->os[2];  // T2
    if (p->ndim == 2) return blocks * p->shape[1] * p->os[2];             // T1
    return 0;
}

static int launch_x(const pf_plan* p, const pf_batch& b, bool fwd, const float2* src, float2* dst, float factor,
                    hipStream_t s) {
    const pf_fftk_axis& X = p->fftk->ax[2];
    const int KCv = b.K * b.Cv;
    if (X.G == 256 && b.L == 1 && !b.b0s) {
        const int BX = fwd ? X256_BF : X256_BA;
        const int KC = std::min(KCv, BX);
        const int YG = BX / KC;
        const int nkc = (KCv + KC - 1) / KC;
        const int64_t rows = (int64_t)b.nfg * b.Cb * p->shape[0] * p->shape[1];
        const int64_t nblk = ((rows + YG - 1) / YG) * nkc;
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
        Geo g = geo_of(p);
        Batch t = batch_of(b);
        if (fwd && KC < 8)
            hipLaunchKernelGGL((k_fftx256<true, X256_BF, true>), dim3((unsigned)nblk), dim3(16 * X256_BF), 0, s, g,
                               t, X.tw, p->apod, src, dst, KC, YG, nkc, 1.0f);
        else if (fwd)
            hipLaunchKernelGGL((k_fftx256<true, X256_BF>), dim3((unsigned)nblk), dim3(16 * X256_BF), 0, s, g, t,
                               X.tw, p->apod, src, dst, KC, YG, nkc, 1.0f);
        else
            hipLaunchKernelGGL((k_fftx256<false, X256_BA>), dim3((unsigned)nblk), dim3(16 * X256_BA), 0, s, g, t,
                               X.tw, p->apod, src, dst, KC, YG, nkc, factor);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    const bool x512inv = X.G == 512 && PF_FFT512X32 && !fwd && b.L == 1 && !b.b0s;
    const bool t8 = PF_FFTQ && colq8_supported(X.G);  // 8 threads per line: 32 columns per block
    if (PF_FFTQ && (colq_supported(X.G) || t8 || x512inv) && b.L <= X256_BF) {
        const int BX = t8 ? 2 * X256_BF : X256_BF;  // columns per block, both directions
        const bool seg = !fwd && (b.L > 1 || b.b0s);  // inverse summing field-map segments
        int KC, YG, nkc;
        if (seg) {  // all L segments of KC pairs (of YG rows when they fit) per block
            KC = std::min(KCv, std::max(1, BX / b.L));
            YG = KC == KCv ? std::max(1, BX / (b.L * KCv)) : 1;
            nkc = (KCv + KC - 1) / KC;
        } else {  // forward columns v = l*KCv + kc over all segments
            const int VT = fwd ? b.L * KCv : KCv;
            KC = std::min(VT, BX);
            YG = BX / KC;
            nkc = (VT + KC - 1) / KC;
        }
        const int64_t rows = (int64_t)b.nfg * b.Cb * p->shape[0] * p->shape[1];
        const int64_t nblk = ((rows + YG - 1) / YG) * nkc;
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
        if (nblk < 1) return PF_OK;
        Geo g = geo_of(p);
        Batch t = batch_of(b);
        const dim3 grid((unsigned)nblk), blk(16 * X256_BF);
        if (t8) {
#define PF_XQ8_CASE(QV)                                                                                         \
    case QV:                                                                                                    \
        if (fwd && KC < 8)                                                                                      \
            hipLaunchKernelGGL((k_fftxq<true, QV, 2 * X256_BF, true, false, 8>), grid, blk, 0, s, g, t, X.tw,    \
                               p->apod, src, dst, KC, YG, nkc, 1.0f);                                           \
        else if (fwd)                                                                                           \
            hipLaunchKernelGGL((k_fftxq<true, QV, 2 * X256_BF, false, false, 8>), grid, blk, 0, s, g, t, X.tw,   \
                               p->apod, src, dst, KC, YG, nkc, 1.0f);                                           \
        else if (seg)                                                                                           \
            hipLaunchKernelGGL((k_fftxq<false, QV, 2 * X256_BF, false, true, 8>), grid, blk, 0, s, g, t, X.tw,   \
                               p->apod, src, dst, KC, YG, nkc, factor);                                         \
        else                                                                                                    \
            hipLaunchKernelGGL((k_fftxq<false, QV, 2 * X256_BF, false, false, 8>), grid, blk, 0, s, g, t, X.tw,  \
                               p->apod, src, dst, KC, YG, nkc, factor);                                         \
        break;
            switch (X.G / 8) {
                PF_XQ8_CASE(5)
                PF_XQ8_CASE(7)
                PF_XQ8_CASE(9)
                PF_XQ8_CASE(15)
                PF_XQ8_CASE(21)
                PF_XQ8_CASE(25)
                PF_XQ8_CASE(35)
                PF_XQ8_CASE(45)
                default:
                    return pf_fail(PF_ERR_FFT, "no register row pass for this size");
            }
#undef PF_XQ8_CASE
            PF_HIP(hipGetLastError());
            return PF_OK;
        }
        if (x512inv) {  // inverse only: the forward's 32-value form measured slower than k_fftx512_fwd
            hipLaunchKernelGGL((k_fftxq<false, 32, X256_BF>), grid, blk, 0, s, g, t, X.tw, p->apod, src, dst, KC, YG,
                               nkc, factor);
            PF_HIP(hipGetLastError());
            return PF_OK;
        }
#define PF_XQ_CASE(QV)                                                                                          \
    case QV:                                                                                                    \
        if (fwd && KC < 8)                                                                                      \
            hipLaunchKernelGGL((k_fftxq<true, QV, X256_BF, true>), grid, blk, 0, s, g, t, X.tw, p->apod, src,     \
                               dst, KC, YG, nkc, 1.0f);                                                         \
        else if (fwd)                                                                                           \
            hipLaunchKernelGGL((k_fftxq<true, QV, X256_BF>), grid, blk, 0, s, g, t, X.tw, p->apod, src, dst, KC,  \
                               YG, nkc, 1.0f);                                                                  \
        else if (seg)                                                                                           \
            hipLaunchKernelGGL((k_fftxq<false, QV, X256_BF, false, true>), grid, blk, 0, s, g, t, X.tw, p->apod,  \
                               src, dst, KC, YG, nkc, factor);                                                  \
        else                                                                                                    \
            hipLaunchKernelGGL((k_fftxq<false, QV, X256_BF>), grid, blk, 0, s, g, t, X.tw, p->apod, src, dst, KC, \
                               YG, nkc, factor);                                                                \
        break;
        switch (X.G / 16) {
            PF_XQ_CASE(5)
            PF_XQ_CASE(6)
            PF_XQ_CASE(7)
            PF_XQ_CASE(8)
            PF_XQ_CASE(10)
            PF_XQ_CASE(12)
            PF_XQ_CASE(14)
            PF_XQ_CASE(15)
            PF_XQ_CASE(20)
            PF_XQ_CASE(21)
            PF_XQ_CASE(24)
            PF_XQ_CASE(9)
            PF_XQ_CASE(18)
            PF_XQ_CASE(25)
            PF_XQ_CASE(28)
            PF_XQ_CASE(30)
            PF_XQ_CASE(36)
            PF_XQ_CASE(40)
            default:
                return pf_fail(PF_ERR_FFT, "no register row pass for this size");
        }
#undef PF_XQ_CASE
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (X.G == 512 && PF_FFT512 && b.L <= X512_B) {
        Geo g = geo_of(p);
        Batch t = batch_of(b);
        const int64_t rows = (int64_t)b.nfg * b.Cb * p->shape[0] * p->shape[1];
        if (fwd) {
            const int64_t nblk = (rows * b.R * b.Cv + X512_B - 1) / X512_B;
            if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
            if (nblk < 1) return PF_OK;
            if (b.R * b.Cv <= 2)
                hipLaunchKernelGGL(k_fftx512_fwd<true>, dim3((unsigned)nblk), dim3(32 * X512_B), 0, s, g, t, X.tw,
                                   p->apod, src, dst);
            else
                hipLaunchKernelGGL(k_fftx512_fwd<false>, dim3((unsigned)nblk), dim3(32 * X512_B), 0, s, g, t, X.tw,
                                   p->apod, src, dst);
        } else {
            XBlk xb;  // all L segments of KC (k, cv) pairs per block, YG rows when they fit
            xb.KC = std::min(KCv, std::max(1, X512_B / b.L));
            xb.YG = xb.KC == KCv ? std::max(1, X512_B / (b.L * KCv)) : 1;
            xb.nkc = (KCv + xb.KC - 1) / xb.KC;
            const int64_t nblk = ((rows + xb.YG - 1) / xb.YG) * xb.nkc;
            if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
            if (nblk < 1) return PF_OK;
            hipLaunchKernelGGL(k_fftx512_adj, dim3((unsigned)nblk), dim3(32 * X512_B), 0, s, g, t, X.tw, p->apod,
                               src, dst, xb, factor);
        }
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (X.rad.empty())
        return pf_fail(PF_ERR_UNSUPPORTED, kNoStockham);
    AxisFFT A = axis_desc(X);
    if ((int64_t)b.L * X.G > GBMAX)
        return pf_fail(PF_ERR_UNSUPPORTED, "too many B0 segments for the fused FFT at this grid size");
    XBlk xb;
    const int cols = std::max(b.L, cols_for(X.G));  // columns per block, >= L
    xb.KC = std::max(1, std::min(KCv, cols / b.L));
    xb.YG = std::max(1, cols / (b.L * xb.KC));
    while (xb.YG > 1 && (int64_t)xb.YG * b.L * xb.KC * X.G > GBMAX) xb.YG /= 2;
    while (xb.KC > 1 && (int64_t)xb.YG * b.L * xb.KC * X.G > GBMAX) xb.KC /= 2;
    xb.nkc = (KCv + xb.KC - 1) / xb.KC;
    const int64_t rows = (int64_t)b.nfg * b.Cb * p->shape[0] * p->shape[1];
    const int64_t nblk = ((rows + xb.YG - 1) / xb.YG) * xb.nkc;
    if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
    const int B = xb.YG * b.L * xb.KC;
    size_t lds = ((size_t)X.G * (B + 1) + X.G) * sizeof(float2);
    Geo g = geo_of(p);
    Batch t = batch_of(b);
    if (fwd)
        hipLaunchKernelGGL(k_fftx_fwd, dim3((unsigned)nblk), dim3(FT), lds, s, g, t, A, p->apod, src, dst, xb);
    else
        hipLaunchKernelGGL(k_fftx_adj, dim3((unsigned)nblk), dim3(FT), lds, s, g, t, A, p->apod, src, dst, xb,
                           factor);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

static int launch_col(const pf_plan* p, int axis, bool fwd, int64_t O, int64_t I, const float2* src, float2* dst,
                      hipStream_t s, ColMask cm = ColMask{nullptr, 0, 1}) {
    const pf_fftk_axis& X = p->fftk->ax[axis];
    if (X.G == 512 && PF_FFT512) {
        const bool wide = O * I >= C512_MIN;
        const int B = wide ? C512_B : C512_BS;
        const int64_t nblk = O * ((I + B - 1) / B);
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
        if (nblk < 1) return PF_OK;
        if (wide && fwd)
            hipLaunchKernelGGL((k_fftcol512<true, C512_B>), dim3((unsigned)nblk), dim3(16 * C512_B), 0, s, X.N, I,
                               X.tw, src, dst, cm);
        else if (wide)
            hipLaunchKernelGGL((k_fftcol512<false, C512_B>), dim3((unsigned)nblk), dim3(16 * C512_B), 0, s, X.N, I,
                               X.tw, src, dst, cm);
        else if (fwd)
            hipLaunchKernelGGL((k_fftcol512_t32<true, C512_BS>), dim3((unsigned)nblk), dim3(32 * C512_BS), 0, s,
                               X.N, I, X.tw, src, dst, cm);
        else
            hipLaunchKernelGGL((k_fftcol512_t32<false, C512_BS>), dim3((unsigned)nblk), dim3(32 * C512_BS), 0, s,
                               X.N, I, X.tw, src, dst, cm);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (X.G == 256) {
        const int B = fwd ? C256_BF : C256_BA;
        const int64_t nblk = O * ((I + B - 1) / B);
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
        if (nblk < 1) return PF_OK;
        if (fwd)
            hipLaunchKernelGGL((k_fftcol256<true, C256_BF>), dim3((unsigned)nblk), dim3(16 * C256_BF), 0, s, X.N, I,
                               X.tw, src, dst, cm);
        else
            hipLaunchKernelGGL((k_fftcol256<false, C256_BA>), dim3((unsigned)nblk), dim3(16 * C256_BA), 0, s, X.N, I,
                               X.tw, src, dst, cm);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (PF_FFTQ && colq8_supported(X.G)) {  // 8 threads x Q values per column, 256-thread blocks
        const int64_t nblk = O * ((I + CQ_B - 1) / CQ_B);
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
        if (nblk < 1) return PF_OK;
        const dim3 grid((unsigned)nblk), blk(8 * CQ_B);
#define PF_COLQ8_CASE(QV)                                                                                     \
    case QV:                                                                                                 \
        if (fwd)                                                                                             \
            hipLaunchKernelGGL((k_fftcolq<true, QV, CQ_B, 8>), grid, blk, 0, s, X.N, I, X.tw, src, dst, cm); \
        else                                                                                                 \
            hipLaunchKernelGGL((k_fftcolq<false, QV, CQ_B, 8>), grid, blk, 0, s, X.N, I, X.tw, src, dst, cm); \
        break;
        switch (X.G / 8) {
            PF_COLQ8_CASE(5)
            PF_COLQ8_CASE(7)
            PF_COLQ8_CASE(9)
            PF_COLQ8_CASE(15)
            PF_COLQ8_CASE(21)
            PF_COLQ8_CASE(25)
            PF_COLQ8_CASE(35)
            PF_COLQ8_CASE(45)
            default:
                return pf_fail(PF_ERR_FFT, "no register column pass for this size");
        }
#undef PF_COLQ8_CASE
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (PF_FFTQ && colq_supported(X.G)) {
        const int64_t nblk = O * ((I + CQ_B - 1) / CQ_B);
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
        if (nblk < 1) return PF_OK;
        const dim3 grid((unsigned)nblk), blk(16 * CQ_B);
        const int Q = X.G / 16;
#define PF_COLQ_CASE(QV)                                                                                     \
    case QV:                                                                                                 \
        if (fwd)                                                                                             \
            hipLaunchKernelGGL((k_fftcolq<true, QV, CQ_B>), grid, blk, 0, s, X.N, I, X.tw, src, dst, cm);   \
        else                                                                                                 \
            hipLaunchKernelGGL((k_fftcolq<false, QV, CQ_B>), grid, blk, 0, s, X.N, I, X.tw, src, dst, cm);  \
        break;
        switch (Q) {
            PF_COLQ_CASE(5)
            PF_COLQ_CASE(6)
            PF_COLQ_CASE(7)
            PF_COLQ_CASE(8)
            PF_COLQ_CASE(10)
            PF_COLQ_CASE(12)
            PF_COLQ_CASE(14)
            PF_COLQ_CASE(15)
            PF_COLQ_CASE(20)
            PF_COLQ_CASE(21)
            PF_COLQ_CASE(24)
            PF_COLQ_CASE(9)
            PF_COLQ_CASE(18)
            PF_COLQ_CASE(25)
            PF_COLQ_CASE(28)
            PF_COLQ_CASE(30)
            PF_COLQ_CASE(36)
            PF_COLQ_CASE(40)
            default:
                return pf_fail(PF_ERR_FFT, "no register column pass for this size");
        }
#undef PF_COLQ_CASE
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (cm.mode != 0) return pf_fail(PF_ERR_UNSUPPORTED, "column mask on an LDS FFT pass");
    if (X.rad.empty())
        return pf_fail(PF_ERR_UNSUPPORTED, kNoStockham);
    AxisFFT A = axis_desc(X);
    const int B = cols_for(X.G);
    int bshift = 0;
    while ((1 << bshift) < B) ++bshift;
    const int64_t nblk = O * ((I + B - 1) / B);
    if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
    if (nblk < 1) return PF_OK;
    size_t lds = ((size_t)X.G * (B + 1) + X.G) * sizeof(float2);
    if (fwd)
        hipLaunchKernelGGL(k_fftcol<true>, dim3((unsigned)nblk), dim3(FT), lds, s, A, I, B, bshift, src, dst);
    else
        hipLaunchKernelGGL(k_fftcol<false>, dim3((unsigned)nblk), dim3(FT), lds, s, A, I, B, bshift, src, dst);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

// the inverse z pass of a 3-D plan with a column mask runs a register
// kernel, so the fused adjoint may leave the dead columns unwritten
bool pf_fftk_masked(const pf_plan* p) {
    if (!PF_COLMASK || p->ndim != 3 || !p->colmask || !p->fftk) return false;
    const int G = p->fftk->ax[0].G;
    return G == 256 || (G == 512 && PF_FFT512) || (PF_FFTQ && (colq_supported(G) || colq8_supported(G)));
}

// image -> grid (deapodise, pad, FFT); `extra` holds pf_fftk_extra_bytes
int pf_fftk_forward(pf_plan* p, const pf_batch& b, const float2* img, float2* grid, float2* extra, hipStream_t s) {
    int rc = pf_fftk_ready(p);
    if (rc != PF_OK) return rc;
    const int64_t RC = (int64_t)b.R * b.Cv;
    const int64_t FC = (int64_t)b.nfg * b.Cb;
    if (p->ndim == 1) return launch_x(p, b, true, img, grid, 1.0f, s);
    if (p->ndim == 2) {
        if ((rc = launch_x(p, b, true, img, extra, 1.0f, s)) != PF_OK) return rc;
        return launch_col(p, 1, true, FC, (int64_t)p->os[2] * RC, extra, grid, s);
    }
    // 3-D: T1 lives in the grid buffer (dead until the Z pass writes it)
    float2* t1 = grid;
    float2* t2 = extra;
    if ((rc = launch_x(p, b, true, img, t1, 1.0f, s)) != PF_OK) return rc;
    if ((rc = launch_col(p, 1, true, FC * p->shape[0], (int64_t)p->os[2] * RC, t1, t2, s)) != PF_OK) return rc;
    return launch_col(p, 0, true, FC, (int64_t)p->os[1] * p->os[2] * RC, t2, grid, s);
}

// grid -> image (inverse FFT, crop, deapodise, times factor); the grid is
// consumed; masked: the z pass takes the dead (y, x) columns as zero
// without reading them (the fused adjoint's spread leaves them unwritten)
int pf_fftk_adjoint(pf_plan* p, const pf_batch& b, float2* grid, float2* img, float2* extra, float factor,
                    hipStream_t s, bool masked) {
    int rc = pf_fftk_ready(p);
    if (rc != PF_OK) return rc;
    const int64_t RC = (int64_t)b.R * b.Cv;
    const int64_t FC = (int64_t)b.nfg * b.Cb;
    if (p->ndim == 1) return launch_x(p, b, false, grid, img, factor, s);
    if (p->ndim == 2) {
        if ((rc = launch_col(p, 1, false, FC, (int64_t)p->os[2] * RC, grid, extra, s)) != PF_OK) return rc;
        return launch_x(p, b, false, extra, img, factor, s);
    }
    float2* t2 = extra;
    float2* t1 = grid;
    const ColMask cz{p->colmask, masked && pf_fftk_masked(p) ? 1 : 0, (int)RC};
    if ((rc = launch_col(p, 0, false, FC, (int64_t)p->os[1] * p->os[2] * RC, grid, t2, s, cz)) != PF_OK) return rc;
    if ((rc = launch_col(p, 1, false, FC * p->shape[0], (int64_t)p->os[2] * RC, t2, t1, s)) != PF_OK) return rc;
    return launch_x(p, b, false, t1, img, factor, s);
}

// Masked centred FFT on the caller's layout (planes, Nz, Ny, Nx), in place
// after the first pass; see the Cartesian pass kernels above.
static int cart_col(const pf_plan* p, int axis, bool fwd, int64_t O, int64_t I, const float2* src, float2* dst,
                    const CartEpi& e, hipStream_t s) {
    const pf_fftk_axis& X = p->fftk->ax[axis];
    if (X.G == 256) {
        const int B = fwd ? CC256_BF : CC256_BA;
        const int64_t nblk = O * ((I + B - 1) / B);
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
        if (nblk < 1) return PF_OK;
        if (fwd)
            hipLaunchKernelGGL((k_cartcol256<true, CC256_BF>), dim3((unsigned)nblk), dim3(16 * CC256_BF), 0, s, I,
                               X.tw, src, dst, e);
        else
            hipLaunchKernelGGL((k_cartcol256<false, CC256_BA>), dim3((unsigned)nblk), dim3(16 * CC256_BA), 0, s, I,
                               X.tw, src, dst, e);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (PF_FFTQ && colq8_supported(X.G)) {  // 8 threads per line, 32 columns per block
        constexpr int B = 2 * CC256_BF;
        const int64_t nblk = O * ((I + B - 1) / B);
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
        if (nblk < 1) return PF_OK;
        const dim3 grid((unsigned)nblk), blk(8 * B);
#define PF_CCQ8_CASE(QV)                                                                                      \
    case QV:                                                                                                  \
        if (fwd)                                                                                              \
            hipLaunchKernelGGL((k_cartcolq<true, QV, B, 8>), grid, blk, 0, s, I, X.tw, src, dst, e);         \
        else                                                                                                  \
            hipLaunchKernelGGL((k_cartcolq<false, QV, B, 8>), grid, blk, 0, s, I, X.tw, src, dst, e);        \
        break;
        switch (X.G / 8) {
            PF_CCQ8_CASE(5)
            PF_CCQ8_CASE(7)
            PF_CCQ8_CASE(9)
            PF_CCQ8_CASE(15)
            PF_CCQ8_CASE(21)
            PF_CCQ8_CASE(25)
            PF_CCQ8_CASE(35)
            PF_CCQ8_CASE(45)
            default:
                return pf_fail(PF_ERR_FFT, "no register Cartesian column pass for this size");
        }
#undef PF_CCQ8_CASE
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (PF_FFTQ && (colq_supported(X.G) || X.G == 512)) {
        constexpr int B = CC256_BF;
        const int64_t nblk = O * ((I + B - 1) / B);
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
        if (nblk < 1) return PF_OK;
        const dim3 grid((unsigned)nblk), blk(16 * B);
#define PF_CCQ_CASE(QV)                                                                                       \
    case QV:                                                                                                  \
        if (fwd)                                                                                              \
            hipLaunchKernelGGL((k_cartcolq<true, QV, B>), grid, blk, 0, s, I, X.tw, src, dst, e);            \
        else                                                                                                  \
            hipLaunchKernelGGL((k_cartcolq<false, QV, B>), grid, blk, 0, s, I, X.tw, src, dst, e);           \
        break;
        switch (X.G / 16) {
            PF_CCQ_CASE(5)
            PF_CCQ_CASE(6)
            PF_CCQ_CASE(7)
            PF_CCQ_CASE(8)
            PF_CCQ_CASE(9)
            PF_CCQ_CASE(10)
            PF_CCQ_CASE(12)
            PF_CCQ_CASE(14)
            PF_CCQ_CASE(15)
            PF_CCQ_CASE(18)
            PF_CCQ_CASE(20)
            PF_CCQ_CASE(21)
            PF_CCQ_CASE(24)
            PF_CCQ_CASE(25)
            PF_CCQ_CASE(28)
            PF_CCQ_CASE(30)
            PF_CCQ_CASE(32)
            PF_CCQ_CASE(36)
            PF_CCQ_CASE(40)
            default:
                return pf_fail(PF_ERR_FFT, "no register Cartesian column pass for this size");
        }
#undef PF_CCQ_CASE
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (X.rad.empty())
        return pf_fail(PF_ERR_UNSUPPORTED, kNoStockham);
    AxisFFT A = axis_desc(X);
    const int B = cols_for(X.G);
    int bshift = 0;
    while ((1 << bshift) < B) ++bshift;
    const int64_t nblk = O * ((I + B - 1) / B);
    if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT columns");
    if (nblk < 1) return PF_OK;
    const size_t lds = ((size_t)X.G * (B + 1) + X.G) * sizeof(float2);
    if (lds > 64 * 1024) {
        static bool raised = false;
        if (!raised) {
            const void* k[] = {reinterpret_cast<const void*>(k_cartcol<true>),
                               reinterpret_cast<const void*>(k_cartcol<false>)};
            for (const void* f : k)
                if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
                    return pf_fail(PF_ERR_HIP, "could not raise the Cartesian FFT kernels' LDS limit");
            raised = true;
        }
    }
    if (fwd)
        hipLaunchKernelGGL(k_cartcol<true>, dim3((unsigned)nblk), dim3(FT), lds, s, A, I, B, bshift, src, dst, e);
    else
        hipLaunchKernelGGL(k_cartcol<false>, dim3((unsigned)nblk), dim3(FT), lds, s, A, I, B, bshift, src, dst, e);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

static int cart_row(const pf_plan* p, bool fwd, int64_t nrows, const float2* src, float2* dst, const CartEpi& e,
                    hipStream_t s) {
    const pf_fftk_axis& X = p->fftk->ax[2];
    if (X.G == 256) {
        const int64_t nblk = (nrows + 15) / 16;
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
        if (nblk < 1) return PF_OK;
        if (fwd)
            hipLaunchKernelGGL(k_cartrow256<true>, dim3((unsigned)nblk), dim3(256), 0, s, nrows, X.tw, src, dst, e);
        else
            hipLaunchKernelGGL(k_cartrow256<false>, dim3((unsigned)nblk), dim3(256), 0, s, nrows, X.tw, src, dst, e);
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (PF_FFTQ && colq8_supported(X.G)) {  // 8 threads per line, 32 rows per block
        const int64_t nblk = (nrows + 31) / 32;
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
        if (nblk < 1) return PF_OK;
        const dim3 grid((unsigned)nblk), blk(256);
#define PF_CRQ8_CASE(QV)                                                                                      \
    case QV:                                                                                                  \
        if (fwd)                                                                                              \
            hipLaunchKernelGGL((k_cartrowq<true, QV, 8>), grid, blk, 0, s, nrows, X.tw, src, dst, e);        \
        else                                                                                                  \
            hipLaunchKernelGGL((k_cartrowq<false, QV, 8>), grid, blk, 0, s, nrows, X.tw, src, dst, e);       \
        break;
        switch (X.G / 8) {
            PF_CRQ8_CASE(5)
            PF_CRQ8_CASE(7)
            PF_CRQ8_CASE(9)
            PF_CRQ8_CASE(15)
            PF_CRQ8_CASE(21)
            PF_CRQ8_CASE(25)
            PF_CRQ8_CASE(35)
            PF_CRQ8_CASE(45)
            default:
                return pf_fail(PF_ERR_FFT, "no register Cartesian row pass for this size");
        }
#undef PF_CRQ8_CASE
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (PF_FFTQ && (colq_supported(X.G) || X.G == 512)) {
        const int64_t nblk = (nrows + 15) / 16;
        if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
        if (nblk < 1) return PF_OK;
        const dim3 grid((unsigned)nblk), blk(256);
#define PF_CRQ_CASE(QV)                                                                                       \
    case QV:                                                                                                  \
        if (fwd)                                                                                              \
            hipLaunchKernelGGL((k_cartrowq<true, QV>), grid, blk, 0, s, nrows, X.tw, src, dst, e);           \
        else                                                                                                  \
            hipLaunchKernelGGL((k_cartrowq<false, QV>), grid, blk, 0, s, nrows, X.tw, src, dst, e);          \
        break;
        switch (X.G / 16) {
            PF_CRQ_CASE(5)
            PF_CRQ_CASE(6)
            PF_CRQ_CASE(7)
            PF_CRQ_CASE(8)
            PF_CRQ_CASE(9)
            PF_CRQ_CASE(10)
            PF_CRQ_CASE(12)
            PF_CRQ_CASE(14)
            PF_CRQ_CASE(15)
            PF_CRQ_CASE(18)
            PF_CRQ_CASE(20)
            PF_CRQ_CASE(21)
            PF_CRQ_CASE(24)
            PF_CRQ_CASE(25)
            PF_CRQ_CASE(28)
            PF_CRQ_CASE(30)
            PF_CRQ_CASE(32)
            PF_CRQ_CASE(36)
            PF_CRQ_CASE(40)
            default:
                return pf_fail(PF_ERR_FFT, "no register Cartesian row pass for this size");
        }
#undef PF_CRQ_CASE
        PF_HIP(hipGetLastError());
        return PF_OK;
    }
    if (X.rad.empty())
        return pf_fail(PF_ERR_UNSUPPORTED, kNoStockham);
    AxisFFT A = axis_desc(X);
    const int B = cols_for(X.G);
    const int64_t nblk = (nrows + B - 1) / B;
    if (nblk > 0x7fffffff) return pf_fail(PF_ERR_UNSUPPORTED, "too many FFT rows");
    if (nblk < 1) return PF_OK;
    const size_t lds = ((size_t)X.G * (B + 1) + X.G) * sizeof(float2);
    if (lds > 64 * 1024) {
        static bool raised = false;
        if (!raised) {
            const void* k[] = {reinterpret_cast<const void*>(k_cartrow<true>),
                               reinterpret_cast<const void*>(k_cartrow<false>)};
            for (const void* f : k)
                if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
                    return pf_fail(PF_ERR_HIP, "could not raise the Cartesian FFT kernels' LDS limit");
            raised = true;
        }
    }
    if (fwd)
        hipLaunchKernelGGL(k_cartrow<true>, dim3((unsigned)nblk), dim3(FT), lds, s, A, nrows, B, src, dst, e);
    else
        hipLaunchKernelGGL(k_cartrow<false>, dim3((unsigned)nblk), dim3(FT), lds, s, A, nrows, B, src, dst, e);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_fftk_cart(pf_plan* p, int64_t planes, int T, const float* mask, int mframes, bool fwd, const float2* src,
                 float2* dst, float factor, hipStream_t s) {
    int rc = pf_fftk_ready(p);
    if (rc != PF_OK) return rc;
    const int64_t Nz = p->shape[0], Ny = p->shape[1], Nx = p->shape[2];
    const CartEpi none{nullptr, 1, 0, 1.0f};
    const CartEpi last{mask, T, mframes, factor};              // forward: last pass
    const CartEpi first{mask, T, mframes, 1.0f};               // inverse: first pass
    const CartEpi fin{nullptr, 1, 0, factor};                   // inverse: last pass
    if (p->ndim == 1) {
        const CartEpi one{mask, T, mframes, factor};
        return cart_row(p, fwd, planes, src, dst, one, s);
    }
    if (fwd) {
        if ((rc = cart_row(p, true, planes * Nz * Ny, src, dst, none, s)) != PF_OK) return rc;
        if (p->ndim == 2) return cart_col(p, 1, true, planes, Nx, dst, dst, last, s);
        if ((rc = cart_col(p, 1, true, planes * Nz, Nx, dst, dst, none, s)) != PF_OK) return rc;
        return cart_col(p, 0, true, planes, Ny * Nx, dst, dst, last, s);
    }
    if (p->ndim == 2) {
        if ((rc = cart_col(p, 1, false, planes, Nx, src, dst, first, s)) != PF_OK) return rc;
        return cart_row(p, false, planes * Ny, dst, dst, fin, s);
    }
    if ((rc = cart_col(p, 0, false, planes, Ny * Nx, src, dst, first, s)) != PF_OK) return rc;
    if ((rc = cart_col(p, 1, false, planes * Nz, Nx, dst, dst, none, s)) != PF_OK) return rc;
    return cart_row(p, false, planes * Nz * Ny, dst, dst, fin, s);
}
