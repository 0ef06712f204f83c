// Internal declarations of the pyfourier_amd NUFFT engine (gfx950 / CDNA4).
//
// Data layout in HBM (all complex data are interleaved float2 = complex64):
//
//   oversampled grid   (nfg, Cb, Gz, Gy, Gx, R, Cv)
//       nfg  frame grids: F for the plain multi-frame operator, 1 when a
//            subspace basis maps every frame onto the same K coefficient grids
//       Cb   coil blocks, C = Cb*Cv batch items (coils, slices, ...)
//       G*   grid points in *stored* (un-shifted) FFT order: the centred
//            index j of the reference is stored at (j - G//2) mod G, so the
//            fftshift/ifftshift pairs of the reference fold into index maps
//       R    reduction vector: L B0 segments x K subspace coefficients
//       Cv   coils innermost, so one tap of one sample reads Cv*R contiguous
//            complex values
//   samples (sorted)   (Mtot, C) in tile-sorted order (coil innermost)
//   samples (user)     (C, F, npts) complex64, the reference's layout
//
// Every 1-D/2-D problem is stored as 3-D with leading unit axes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pfhip.h"

#define PF_WMAX 8

struct pf_fft_cache;

// Spread work list: for every (frame grid, tile) the sorted samples whose
// footprint touches the tile, split into chunks of at most max_chunk entries.
// Tiles with several chunks accumulate into partial tiles that a second
// kernel sums in chunk order (deterministic).
// spread work items over one sorted entry list
struct pf_sitems {
    int4* sitems = nullptr;      // (key, e0, e1, slot or -1)
    int nsitems = 0;
    // the same items without the empty tiles whose (y, x) columns no sample
    // covers (3-D plans with a column mask; the fused adjoint, whose inverse
    // FFT never reads those columns); null when there is no mask
    int4* lsitems = nullptr;
    int nlsitems = 0;
    int4* ritems = nullptr;      // (key, slot0, nslot, 0)
    int nritems = 0;
    int nslots = 0;
};

constexpr int PF_SLIST_PAD = 16;  // readable ids past the end of a spread list

struct pf_slist {
    int32_t* slist = nullptr;    // sorted sample ids grouped by key
    int64_t nslist = 0;
    // [0] heavy tiles split; [1] split further so that the items alone fill
    // the GPU; [2] as [1] with twice the item size, for batches of several
    // value chunks (each item is then that many waves).  Arrays are shared
    // where a split changes nothing.
    pf_sitems split[3];
};

// waves of a spread launch above which the coarser split is used: a batch
// with several value chunks per tile fills the GPU without splitting tiles
// further, and every extra split costs a partial tile per value chunk
#ifndef PF_SPREAD_FILL
#define PF_SPREAD_FILL 8192
#endif
#ifndef PF_COLMASK
#define PF_COLMASK 1  // 3-D plans: skip the (y, x) grid columns no sample covers in the fused operators
#endif
#ifndef PF_SPREAD_MINCH
#define PF_SPREAD_MINCH 64  // smallest work item of the fine split (entries)
#endif

static inline const pf_sitems& pf_sitems_for(const pf_slist& L, int nvch) {
    if ((int64_t)L.split[0].nsitems * nvch >= PF_SPREAD_FILL) return L.split[0];
    return nvch > 1 ? L.split[2] : L.split[1];
}

// Gather column lists (3-D plans with one kernel width on all axes): the
// sorted samples regrouped by window key = (frame, y column, x column, z
// window) of their first tap, where a column is PF_GCOL x PF_GCOL grid points
// in (y, x) and a z window pf_gather_dz(W) planes; every key's grid block
// (plus the kernel halo) is staged in LDS once per coil pair and all its
// samples read their taps from there.
#ifndef PF_GCOL
#define PF_GCOL 16
#endif
#ifndef PF_GCOL_MAXS
#define PF_GCOL_MAXS 256  // lanes per work item (one sample each, one block)
#endif
// z planes per window: (DZ + W - 1) * (PF_GCOL + W - 1)^2 points of 16 bytes
// within 80 KB (two blocks per CU).  Deeper windows restage less halo per
// staged point but hold more samples, and a window with more samples than a
// block's lanes is staged once per block: measured on 3-D radial at 128^3,
// width 6 is fastest at the full budget (DZ 6, 76 KB; 5 and 4 slower), width
// 4 at DZ 7 (58 KB; the full-budget DZ 11 is 12 % slower).
__host__ __device__ constexpr int pf_gather_dz(int W) {
    return W == 6 ? 6 : (W <= 6 ? 11 - W : (W == 7 ? 3 : 2));
}
static_assert((pf_gather_dz(6) + 5) * 21 * 21 * 16 <= 80 * 1024, "width-6 window within two blocks per CU");
struct pf_glist {
    int32_t* ids = nullptr;  // lane slots: sorted-sample index t or -1, PF_GCOL_MAXS per item
    int4* items = nullptr;   // (key, first lane slot, 0, 0)
    int nitems = 0;
    int ncol[2] = {0, 0};    // columns along y, x
    int nzw = 0;             // z windows
};

struct pf_plan {
    int ndim = 0;          // spatial dims 1..3
    int shape[3];          // image matrix (z, y, x), leading 1s for d<3
    int os[3];             // oversampled grid (z, y, x)
    int width[3];          // kernel width per axis (1 on unit / Cartesian axes)
    float oversamp[3];
    float beta[3];
    double scale = 1.0;    // kernel-sum normalisation (divided out)
    int nframes = 1;
    int64_t npts = 0;      // samples per frame (ragged plans: the total)
    int64_t mtot = 0;      // nframes * npts (ragged plans: the total)
    int tile[3];           // spread/sort tile (z, y, x)
    int ntiles[3];
    int64_t ntile_tot = 0; // tiles per frame
    int rec = 16;          // dwords per sample tap record
    int woff[3];           // offset of each axis' weights inside a record
    int device = 0;

    // device arrays (sorted order unless stated)
    int32_t* perm = nullptr;     // sorted -> original sample
    int32_t* iperm = nullptr;    // original -> sorted
    float* taprec = nullptr;     // (mtot, rec): int start[3], int frame, weights
    // 3-D plans: (Gy, Gx) bytes, 1 where some sample's kernel footprint
    // covers the (y, x) grid column (in any z and frame); the fused adjoint's
    // spread leaves the empty tiles of the other columns unwritten and its
    // inverse z pass takes them as zero (null otherwise)
    uint8_t* colmask = nullptr;
    float* cos_sorted = nullptr; // (mtot, 3) oversampled coordinate (z, y, x)
    float* apod = nullptr;       // deapodisation vectors, Nz + Ny + Nx floats
    int64_t* foff = nullptr;     // ragged plans: frame f owns original samples [foff[f], foff[f+1])
    // spread lists: [0] keyed by (frame, tile) for per-frame grids,
    // [1] keyed by tile only for one grid shared by all frames (subspace);
    // identical (aliased) when nframes == 1
    pf_slist sl[2];
    int max_chunk = 0;    // spread entries per work item (0: automatic)
    pf_glist gl;          // LDS-staged gather lists (empty when not built)

    pf_fft_cache* fft = nullptr;   // rocFFT plans (grid sizes with other prime factors)
    struct pf_fftk* fftk = nullptr; // fused pruned FFT (sizes 2^a 3^b 5^c)
};

// ---- error handling --------------------------------------------------------
void pf_set_error(const std::string& msg);
int pf_fail(int code, const std::string& msg);

#define PF_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return pf_fail(PF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// status codes PF_OK / PF_ERR_* come from the public header

// ---- batch description shared by the operator kernels ---------------------
struct pf_batch {
    int C;        // batch items (coils ...)
    int Cv;       // items innermost per grid point
    int Cb;       // C / Cv
    int R;        // reduction vector length (L*K)
    int L;        // B0 segments (1 = none)
    int K;        // subspace coefficients (1 = none)
    int nfg;      // frame grids (F or 1)
    int nimg;     // image planes per batch item in user layout (F, K or 1)
    const float2* basis;  // (F, K) complex or null
    const float2* b0t;    // (L, nt) temporal interpolator or null
    const float2* b0s;    // (L, Nz*Ny*Nx) spatial maps or null
    int nt;               // readout length for b0t
};

// ---- internal launchers (return PF_OK or an error code) -------------------
int pf_launch_deapod_pad(const pf_plan* p, const pf_batch& b, const float2* img, float2* grid,
                         hipStream_t s);
int pf_launch_crop_deapod(const pf_plan* p, const pf_batch& b, const float2* grid, float2* img,
                          float factor, hipStream_t s);
int pf_launch_gather(const pf_plan* p, const pf_batch& b, const float2* grid, float2* ys,
                     hipStream_t s);
// smax_ready: the sort already wrote the per-sample maxima the matrix-core
// spread scales by (pf_spread_smax), else the spread computes them
// live: only the items of pf_sitems::lsitems (the empty tiles of columns no
// sample covers stay unwritten; for a masked inverse FFT)
int pf_launch_spread(const pf_plan* p, const pf_batch& b, const float2* ys, float2* grid,
                     float2* partial, hipStream_t s, bool smax_ready = false, bool live = false);
// where the matrix-core spread keeps its per-sample maxima inside the spread
// workspace, or null when the batch does not use it
float* pf_spread_smax(const pf_plan* p, const pf_batch& b, float2* partial);
int pf_launch_unsort(const pf_plan* p, const pf_batch& b, const float2* ys, float2* y,
                     const float* weight, float factor, hipStream_t s);
// smax: also write the per-sample maxima of the matrix-core spread where
// the sort can (*smax_written says whether it did)
int pf_launch_sort(const pf_plan* p, const pf_batch& b, const float2* y, float2* ys,
                   const float* weight, hipStream_t s, float* smax = nullptr, bool* smax_written = nullptr);
// The fused adjoint's expanded form of a batch with B0 segments or a basis:
// true (and *bx = the plain batch of R*C items on the same grid memory) when
// all items form one vector (Cb = 1) and that plain batch takes the
// pipelined spread.  pf_launch_sort_expand then writes the sorted samples as
// (Mtot, R*C) with the coefficients applied, for pf_launch_spread(bx).
bool pf_spread_expand(const pf_plan* p, const pf_batch& b, pf_batch* bx);
int pf_launch_sort_expand(const pf_plan* p, const pf_batch& b, const float2* y, float2* ys, const float* weight,
                          hipStream_t s);
int pf_fft_exec(pf_plan* p, const pf_batch& b, float2* grid, int inverse, hipStream_t s);
void pf_fft_destroy(pf_fft_cache* c);
int pf_fft_cache_init(pf_plan* p);  // rocFFT setup + the plan's (empty) cache
size_t pf_spread_partial_bytes(const pf_plan* p, const pf_batch& b);
int pf_spread_coil_chunk(const pf_plan* p, const pf_batch& b);
// names of the gather / spread kernel instances a batch runs (diagnostics)
std::string pf_gather_name(const pf_plan* p, const pf_batch& b);
std::string pf_spread_name(const pf_plan* p, const pf_batch& b);
int pf_ops_prepare();  // kernel attributes, set at plan creation
// the LDS-staged gather applies to this batch (its forward grid is then laid
// out with Cv = 2 coil pairs, Cb = C / 2)
bool pf_gather_col_layout(const pf_plan* p, const pf_batch& b);
int pf_noalloc_prepare();  // this device's pf_nufft_noalloc scratch address, at plan creation

// fused deapodise/pad/FFT and inverse FFT/crop/deapodise (pf_fftk.hip)
bool pf_fftk_supported(const pf_plan* p);
int pf_fftk_init(pf_plan* p);
int pf_fftk_ready(const pf_plan* p);  // PF_OK when pf_fftk_init has run
void pf_fftk_destroy(struct pf_fftk* f);
size_t pf_fftk_extra_bytes(const pf_plan* p, const pf_batch& b);
// masked (the fused adjoint, where pf_fftk_masked(p)): the inverse z pass
// takes the (y, x) columns no sample covers as zero without reading them
bool pf_fftk_masked(const pf_plan* p);
int pf_fftk_forward(pf_plan* p, const pf_batch& b, const float2* img, float2* grid, float2* extra, hipStream_t s);
int pf_fftk_adjoint(pf_plan* p, const pf_batch& b, float2* grid, float2* img, float2* extra, float factor,
                    hipStream_t s, bool masked = false);
// masked centred FFT (oversampling 1) on the layout (planes, Nz, Ny, Nx):
// src -> dst, in place after the first pass; mask (mframes, Nz*Ny*Nx) or null
int pf_fftk_cart(pf_plan* p, int64_t planes, int T, const float* mask, int mframes, bool fwd, const float2* src,
                 float2* dst, float factor, hipStream_t s);

// ---- small helpers --------------------------------------------------------
static inline int64_t pf_gtot(const pf_plan* p) { return (int64_t)p->os[0] * p->os[1] * p->os[2]; }
static inline int64_t pf_ntot(const pf_plan* p) {
    return (int64_t)p->shape[0] * p->shape[1] * p->shape[2];
}
// spread list for a batch: frame-merged when all frames share one grid
static inline const pf_slist& pf_slist_for(const pf_plan* p, const pf_batch& b) {
    return (b.nfg == 1 && p->nframes > 1) ? p->sl[1] : p->sl[0];
}

// XCD-aware block remap: the hardware deals workgroups round-robin over the
// 8 XCDs; remap so each XCD walks a contiguous range of the logical order and
// neighbouring tiles share that XCD's L2.
__device__ __forceinline__ unsigned pf_xcd_remap(unsigned b, unsigned nb) {
    const unsigned nx = 8;
    unsigned per = nb / nx, rem = nb % nx;
    unsigned x = b % nx, i = b / nx;
    // XCDs [0, rem) own per+1 logical blocks, the rest own per
    unsigned base = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    return base + i;
}

// plan helpers shared by the NUFFT and Toeplitz objects
int pf_plan_params(pf_plan* p, int ndim, const int* shape_zyx, const int* width_zyx,
                   const float* oversamp_zyx, int nframes, int64_t npts, const int* tile_zyx, int chunk);
int pf_plan_upload_apod(pf_plan* p);
void pf_plan_free(pf_plan* p);
int pf_plan_build(pf_plan* p, const float* coord, hipStream_t stream);
// register-spread tiles for ndim (3 ints each, z y x); returns how many exist
int pf_spread_reg_tiles(int ndim, int* tiles, int cap);
int pf_launch_gather_dcoord(const pf_plan* p, const pf_batch& b, const float2* grid, const float2* h, const float* scale,
                            bool accumulate, float* out, hipStream_t s);
