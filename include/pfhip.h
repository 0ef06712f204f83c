/* pfhip.h -- C ABI of the pyfourier_amd NUFFT engine (MI355X / gfx950).
 *
 * The reference (INFN-MRI/pyfourier) is pure Python; its plug-in points for
 * this path are Python functions and kernel dispatch tables.  Each entry
 * point below names the reference interface it replaces (paths relative to
 * the reference repository root).
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (hipMalloc / torch CUDA tensors)
 *     unless marked "host".  complex data are interleaved complex64
 *     (float re, float im), passed as void*.
 *   - Every call only enqueues work on `stream` (a hipStream_t passed as
 *     void*, NULL = default stream) and returns without synchronising or
 *     allocating, except the pf_*_create calls, which allocate and
 *     synchronise once when the object is built (fused-FFT twiddle tables
 *     included), so operator calls may be captured into a HIP graph from the
 *     first one.  One exception: grid sizes with prime factors other than
 *     2, 3 and 5 run rocFFT, whose plan and work buffer for a new batch size
 *     are made (under the plan's lock) at the first call with that size.
 *   - Every call returns 0 on success or a PF_ERR_* code; pf_last_error()
 *     returns a thread-local message for the last failure.  Arguments are
 *     validated on the host before any launch.
 *   - A plan is immutable after creation; calls on one plan may be issued
 *     from several threads, but calls that overlap in time on different
 *     streams share the plan's rocFFT work buffer and must be serialised by
 *     the caller.
 *
 * Shapes (reference conventions, src/pyfourier/_nufft/_fwd.py:42-126):
 *   coordinates  (F, npts, ndim) float32 in grid units [-N/2, N/2), last axis
 *                ordered (x, y, z); shape/width/oversamp ordered (z, y, x)
 *   image        (C, nimg, [Nz], Ny, Nx)   nimg = F (plain), K (subspace)
 *   k-space      (C, F, npts)
 */
#ifndef PFHIP_H
#define PFHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_ERR_ARG 1
#define PF_ERR_HIP 2
#define PF_ERR_FFT 3
#define PF_ERR_ALLOC 4
#define PF_ERR_UNSUPPORTED 5

#define PF_NORM_NONE 0  /* forward unscaled, adjoint 1/Ghat (reference norm=None) */
#define PF_NORM_ORTHO 1 /* both 1/sqrt(Ghat) (reference norm="ortho") */
/* The exact adjoint of the *other* direction under PF_NORM_NONE: the forward
 * scaled by 1/Ghat (adjoint of nufft_adj) or the adjoint unscaled (adjoint of
 * nufft).  Used by autograd backward passes (reference defect A.22 fixed). */
#define PF_NORM_ADJOINT 2

typedef struct pf_plan pf_plan;

/* Batch description of one operator call. */
typedef struct pf_batch_desc {
    int batch;                /* C: coils x slices x ... (>= 1) */
    int nseg;                 /* L: B0 time segments, 1 = no field map */
    int ncoeff;               /* K: subspace coefficients, 0 = no basis */
    const void* basis;        /* (F, K) complex64, required when ncoeff > 0 */
    const void* b0_temporal;  /* (L, nt) complex64, required when nseg > 1 */
    const void* b0_spatial;   /* (L, Nz*Ny*Nx) complex64, required when nseg > 1 */
    int nt;                   /* readout length of b0_temporal; the readout
                                 index of sample j of a frame is j % nt */
} pf_batch_desc;

/* Plan summary (host struct). */
typedef struct pf_plan_info {
    int ndim;
    int shape[3];    /* (z, y, x), leading 1 for ndim < 3 */
    int os_shape[3]; /* ceil(oversamp * shape) */
    int width[3];
    float beta[3];
    double scale;    /* kernel-sum normalisation */
    int nframes;
    long long npts;
    int tile[3];
    long long spread_entries; /* (sample, tile) pairs of the spread list */
    int spread_items;         /* work items of the finer split (batches with few value chunks) */
    int spread_slots;         /* partial tiles that split uses */
} pf_plan_info;

/* Message of the last failed call on this thread ("" if none). */
const char* pf_last_error(void);

/* ABI version (major*100 + minor). */
int pf_version(void);

/* Build a plan on the current HIP device.
 * Replaces plan_nufft / Interpolator
 *   src/pyfourier/_nufft/_plan.py:20-199 and
 *   src/pyfourier/_subroutines/_grid/_plan.py:14-114,192-210.
 * coord:   (F*npts, ndim) float32, grid units (already normalised; the
 *          reference normalisation _subroutines/_traj.py:10-54 is done by the
 *          caller)
 * shape, width, oversamp: host arrays of ndim entries, (z, y, x) order; a
 *          Cartesian axis is passed with width 1 and oversamp 1.0
 *          (src/pyfourier/_nufft/_plan.py:139-146). */
int pf_plan_create(const float* coord, int ndim, const int* shape, const int* width,
                   const float* oversamp, int nframes, long long npts, void* stream,
                   pf_plan** out);

/* pf_plan_create with the sort/spread tile chosen by the caller (an
 * autotuner) instead of the built-in default; same reference interface as
 * pf_plan_create, whose plan_nufft has no such knob.
 * tile:  host array of ndim extents (z, y, x order) or NULL for the default;
 *        a tile that fits inside one listed by pf_spread_tiles uses the
 *        register-resident spread, a larger one the LDS-staged spread.
 * chunk: spread-list entries per work item before a tile is split into
 *        chunks that a reduce pass sums in order (0: automatic, 64..1024:
 *        4x the mean tile load, or less so that the list makes ~4096 work
 *        items). */
int pf_plan_create_tiled(const float* coord, int ndim, const int* shape, const int* width,
                         const float* oversamp, int nframes, long long npts, const int* tile,
                         int chunk, void* stream, pf_plan** out);

/* A plan over frames of unequal length ("ragged"): frame f owns samples
 * frame_offsets[f] .. frame_offsets[f+1]-1 of coord (host array of
 * nframes + 1 entries, frame_offsets[0] = 0, non-decreasing).  Each frame is
 * an independent problem of the same image geometry (its own image plane in
 * the (batch, nframes, *shape) layout, its own grid), so an uneven set of
 * slices or trajectories runs in one call; k-space is (batch, total) with the
 * frames concatenated.  Field-map segments are not supported on such plans.
 * No reference counterpart: the reference loops over slices in Python with
 * one plan_nufft (src/pyfourier/_nufft/_plan.py:20-199) per slice.
 * pf_plan_get_info reports npts = total samples. */
int pf_plan_create_ragged(const float* coord, int ndim, const int* shape, const int* width,
                          const float* oversamp, int nframes, const long long* frame_offsets,
                          const int* tile, int chunk, void* stream, pf_plan** out);

/* Tiles (z, y, x) with a register-resident spread for ndim: writes up to
 * `cap` of them as 3 ints each into the host array `tiles`, sets *count to
 * the number available.  No reference counterpart (tuning aid). */
int pf_spread_tiles(int ndim, int* tiles, int cap, int* count);

/* Release a plan and its device buffers (src/pyfourier/_nufft/_plan.py:203-225). */
int pf_plan_destroy(pf_plan* plan);

/* Fill a host pf_plan_info. */
int pf_plan_get_info(const pf_plan* plan, pf_plan_info* info);

/* Tile-sorted order of the plan's samples: copies (F*npts) int32 sorted
 * position -> original sample index into the device buffer `perm`. */
int pf_plan_get_permutation(const pf_plan* plan, int* perm, void* stream);

/* Workspace bytes that pf_nufft (op=0) or pf_nufft_adj (op=1) need for the
 * batch `desc` (device memory the caller allocates; 256-byte aligned). */
int pf_workspace_size(const pf_plan* plan, const pf_batch_desc* desc, int op, size_t* bytes);

/* Forward NUFFT  y = weight * (1/scale) * Gather(FFT(pad(apod * x))).
 * Replaces _nufft_fwd/_do_nufft src/pyfourier/_nufft/_nufft.py:11-90,218-251
 * (with the B0 segment loop, defect A.13 fixed) behind nufft()
 * src/pyfourier/_nufft/_fwd.py:18-230.
 * image (C, nimg, N...) -> kspace (C, F, npts); weight: (F*npts) float32 or
 * NULL; norm: PF_NORM_*. */
int pf_nufft(pf_plan* plan, const pf_batch_desc* desc, const void* image, void* kspace,
             const float* weight, int norm, void* workspace, size_t workspace_bytes, void* stream);

/* Adjoint NUFFT  x = apod * crop(IFFT((1/scale) * Spread(weight * y))).
 * Replaces _nufft_adj/_do_nufft_adj src/pyfourier/_nufft/_nufft.py:93-169,254-287
 * behind nufft_adj() src/pyfourier/_nufft/_adj.py:18-216.  Deterministic:
 * run-to-run bit-identical (no global float atomics). */
int pf_nufft_adj(pf_plan* plan, const pf_batch_desc* desc, const void* kspace, void* image,
                 const float* weight, int norm, void* workspace, size_t workspace_bytes,
                 void* stream);

/* Names of the kernel instances pf_nufft / pf_nufft_adj run for `desc`, as
 * host text "gather=<instance> spread=<instance>[+k_spread_reduce]
 * fft=fused|rocfft" (NUL-terminated, at most cap bytes).  No reference
 * counterpart (the reference's dispatch tables are Python dicts,
 * src/pyfourier/_subroutines/_grid/_cart2noncart/__init__.py:17-18 and
 * _noncart2cart/__init__.py:17-18); lets tests and profiles name what ran. */
int pf_describe_launch(const pf_plan* plan, const pf_batch_desc* desc, char* buf, size_t cap);

/* Fixed workspace of pf_nufft_noalloc: one device array per GPU. */
#define PF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* pf_nufft (op = 0, in = image, out = k-space) or pf_nufft_adj (op = 1,
 * in = k-space, out = image) without a workspace argument and without any
 * allocation: the workspace is a fixed array of PF_NOALLOC_SCRATCH_BYTES
 * bytes in device memory, one per GPU, held until the process ends (its
 * address is looked up when the first plan is created on that GPU).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory at call time (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same results as pf_nufft /
 * pf_nufft_adj.  Limits: the call's pf_workspace_size must not exceed
 * PF_NOALLOC_SCRATCH_BYTES and the grid sizes must be 2^a 3^b 5^c (rocFFT
 * sizes allocate), else PF_ERR_UNSUPPORTED before any launch; the current
 * device must be the plan's.  Every call on one GPU shares its array:
 * calls that overlap in time on one GPU (other streams, other threads)
 * race and must be serialised by the caller. */
int pf_nufft_noalloc(pf_plan* plan, const pf_batch_desc* desc, int op, const void* in, void* out,
                     const float* weight, int norm, void* stream);

/* ---- stage entry points (the reference's kernel plug-in points) ------- */

/* Grid bytes for one call: nfg*C*R*prod(os_shape)*8. */
int pf_grid_size(const pf_plan* plan, const pf_batch_desc* desc, size_t* bytes);

/* Deapodise, zero-pad and ifftshift into the oversampled grid.
 * Replaces _apodize + _resize + ifftshift, src/pyfourier/_nufft/_nufft.py:234-238. */
int pf_deapod_pad(const pf_plan* plan, const pf_batch_desc* desc, const void* image, void* grid,
                  void* stream);

/* Unnormalised in-place FFT of the grid (inverse != 0: e^{+}).
 * Replaces fft/ifft src/pyfourier/_subroutines/_fft.py:11-194 (torch branch). */
int pf_fft(pf_plan* plan, const pf_batch_desc* desc, void* grid, int inverse, void* stream);

/* Gather (forward interpolation) into tile-sorted samples (F*npts, C).
 * Replaces the _degrid dispatch table
 * src/pyfourier/_subroutines/_grid/_cart2noncart/__init__.py:17-18,62-91. */
int pf_interp(const pf_plan* plan, const pf_batch_desc* desc, const void* grid, void* ksp_sorted,
              void* stream);

/* Deterministic spread (adjoint interpolation) of tile-sorted samples into the
 * grid.  Replaces the _grid dispatch table
 * src/pyfourier/_subroutines/_grid/_noncart2cart/__init__.py:17-18,71-100.
 * workspace: pf_spread_workspace_size() bytes (may be NULL when 0). */
int pf_spread(const pf_plan* plan, const pf_batch_desc* desc, const void* ksp_sorted, void* grid,
              void* workspace, void* stream);
int pf_spread_workspace_size(const pf_plan* plan, const pf_batch_desc* desc, size_t* bytes);

/* Byte offsets of the parts of the pf_nufft / pf_nufft_adj workspace:
 * [grid | tile-sorted samples | FFT scratch | spread partial tiles].  The
 * samples part holds (F*npts, C) for the stage entry points; for a batch
 * with B0 segments or a basis whose nseg*ncoeff*batch values per grid point
 * form one vector (batch <= 64) it is sized (F*npts, nseg*ncoeff*batch),
 * which pf_nufft_adj fills with the coefficient-applied values. */
int pf_workspace_offsets(const pf_plan* plan, const pf_batch_desc* desc, size_t* grid_off,
                         size_t* samples_off, size_t* fft_off, size_t* spread_off);

/* Fused deapodise + zero-pad + centred FFT: image (C, nimg, N...) -> grid.
 * For grid sizes 2^a 3^b 5^c, and the 7-smooth sizes 112, 224, 280, 336 and
 * 448 (<= 4096 per axis), a pruned FFT transforms only the lines that hold
 * image data; otherwise pf_deapod_pad + pf_fft.  On the 7-smooth sizes a
 * batch takes at most 16 field-map segments per call.
 * Replaces src/pyfourier/_nufft/_nufft.py:234-241.  fft_workspace: the FFT
 * scratch part of the pf_nufft workspace (pf_workspace_offsets).  Grid
 * layout (F, C/Cv, G..., nseg*ncoeff, Cv), F = 1 with a basis: Cv is the
 * largest divisor of C up to 64, or 2 (coil pairs) for a 3-D plan with one
 * kernel width of 3..8 on all axes and an even C with no field map or basis. */
int pf_pad_fft(pf_plan* plan, const pf_batch_desc* desc, const void* image, void* grid,
               void* fft_workspace, void* stream);

/* Fused inverse centred FFT + crop + deapodise (times `factor`): grid ->
 * image; the grid is overwritten.  Replaces src/pyfourier/_nufft/_nufft.py:275-282. */
int pf_ifft_crop(pf_plan* plan, const pf_batch_desc* desc, void* grid, void* image, float factor,
                 void* fft_workspace, void* stream);

/* Crop, fftshift and deapodise (times `factor`).
 * Replaces _resize + _apodize, src/pyfourier/_nufft/_nufft.py:278-282. */
int pf_crop_deapod(const pf_plan* plan, const pf_batch_desc* desc, const void* grid, void* image,
                   float factor, void* stream);

/* User order (C, F, npts) <-> tile-sorted (F*npts, C), weight fused.
 * Replace the transpose-copies of _cart2noncart/__init__.py:53-54,101-105
 * and _noncart2cart/__init__.py:59-60. */
int pf_sort_samples(const pf_plan* plan, const pf_batch_desc* desc, const void* kspace,
                    void* ksp_sorted, const float* weight, void* stream);
int pf_unsort_samples(const pf_plan* plan, const pf_batch_desc* desc, const void* ksp_sorted,
                      void* kspace, const float* weight, float factor, void* stream);

/* Coordinate derivative of the forward gather (autograd of nufft/nufft_adj
 * with respect to the trajectory).  For tile-sorted h (Mtot, C) and the
 * forward grid (pf_pad_fft of the image), computes per sample
 *   g[a] = Re(sum_c conj(h[s, c]) * d(gather_c)/d(c_os[s, a]))
 * for the oversampled coordinate axes a = (z, y, x), using the analytic
 * derivative of the Kaiser-Bessel tap weights, and writes
 *   dcoord[j, 2 - a] (=|+=) axis_scale[a] * g[a]
 * at the sample's ORIGINAL index j, i.e. dcoord is (Mtot, ndim) float32 in
 * the coordinate order (x, y, z); axis_scale (host, 3 floats, z y x order,
 * NULL = 1) carries 1/scale, the norm factor and d c_os / d coord;
 * accumulate != 0 adds to dcoord (several B0 segment chunks).  Replaces the
 * incomplete _nufft_backward, src/pyfourier/_nufft/_nufft.py:175-209. */
int pf_interp_grad(const pf_plan* plan, const pf_batch_desc* desc, const void* grid, const void* h_sorted,
                   const float* axis_scale, int accumulate, float* dcoord, void* stream);

/* ---- Toeplitz-embedded Gram operator ------------------------------------ */

typedef struct pf_toep pf_toep;

/* Largest per-voxel kernel block R = nseg * ncoeff a Gram operator accepts. */
#define PF_TOEP_RMAX 128

/* Build a Gram operator from its kernel.  Replaces the GramMatrix of
 * plan_spgram, src/pyfourier/_sparsegram/_plan.py:20-270.
 * kernel: (prod(2N), R, R) complex64 device array, R = nseg * ncoeff, the
 *         uncentred FFT of the wrapped point-spread functions on the 2N grid
 *         (index order of an uncentred FFT), already scaled as the operator
 *         needs; copied into the object.  R <= PF_TOEP_RMAX.
 * shape:  host (ndim,) image matrix N (z, y, x).  Embedding sizes 2^a 3^b 5^c
 *         (<= 4096 per axis) run the pruned fused FFT passes, others rocFFT. */
int pf_toeplitz_create(const void* kernel, int ndim, const int* shape, int nseg, int ncoeff, void* stream,
                       pf_toep** out);
int pf_toeplitz_destroy(pf_toep* op);
int pf_toeplitz_workspace_size(const pf_toep* op, int batch, size_t* bytes);

/* y = crop_N(IFFT(T (x) FFT(pad_2N(Ct_l * x)))) summed with conj(Ct_l) over
 * B0 segments; per-voxel R x R kernel for subspace/B0.  Replaces
 * _sparsegram/_do_sparsegram and _bdot, src/pyfourier/_sparsegram/_sparsegram.py:11-81,
 * src/pyfourier/_subroutines/_bmatvec/_bdot.py:16-157.
 * image, out: (batch, ncoeff, N...) complex64; b0_spatial: (nseg, N...) or NULL. */
int pf_toeplitz_apply(pf_toep* op, int batch, const void* image, void* out, const void* b0_spatial,
                      void* workspace, size_t workspace_bytes, void* stream);

/* Dense Cartesian Gram operator: the same object on the N grid itself (no
 * embedding), applied with pf_toeplitz_apply / _workspace_size / _destroy:
 *   y = IFFT(T_f (x) FFT(Ct_l * x)) / prod(N), summed with conj(Ct_l).
 * Replaces plan_gram / GramMatrix and _gram, src/pyfourier/_gram/_plan.py:14-177,
 * src/pyfourier/_gram/_gram.py:11-133.
 * kernel: (nframes, prod(N), R, R) complex64 device array in uncentred FFT
 *         index order, R = nseg * ncoeff: per voxel the matrix
 *         sum_t mask_t * conj(b_l'(x)) * b_l(x) * basis[t, k'] * conj(basis[t, k])
 *         (the mask alone without basis and field map); copied.
 * nframes: kernels per image frame (1 with a basis); the image is then
 *         (batch, nframes, N...), or (batch, ncoeff, N...) with a basis.
 * shape:  host (ndim,) N (z, y, x); sizes with prime factors other than 2, 3
 *         and 5 use rocFFT. */
int pf_gram_create(const void* kernel, int ndim, const int* shape, int nframes, int nseg, int ncoeff,
                   void* stream, pf_toep** out);

/* ---- Cartesian (masked) FFT ---------------------------------------------- */

typedef struct pf_cart pf_cart;

/* Build a Cartesian FFT plan on the current device.  Replaces plan_fft /
 * FFTPlan, src/pyfourier/_fft/_plan.py:14-108 and
 * src/pyfourier/_subroutines/_mask/_plan.py:14-89 (dense branch).
 * mask:  (mframes, Nz*Ny*Nx) float32 device array of sampling weights in
 *        centred k-space order (the reference's binary mask), or NULL for no
 *        mask; copied.  mframes is 1 (shared by every frame) or the frame
 *        count of the calls.
 * shape: host (ndim,) matrix (z, y, x).  Sizes that factor into 2, 3, 5 use
 *        the fused FFT passes, other sizes rocFFT. */
int pf_cart_create(const float* mask, int mframes, int ndim, const int* shape, void* stream, pf_cart** out);
int pf_cart_destroy(pf_cart* plan);

/* Workspace bytes of pf_cart_fft / pf_cart_ifft for `desc` and nframes. */
int pf_cart_workspace_size(const pf_cart* plan, const pf_batch_desc* desc, int nframes, size_t* bytes);

/* kspace_t = factor * mask_t * sum_l b_l(x) sum_k conj(basis[t, k]) FFTc(Ct_l * X_k)
 * Replaces _fft_fwd / _do_fft_fwd / _basis_adj, src/pyfourier/_fft/_fft.py:10-55,101-113,141-151
 * (B0 segments added once, defect A.13 fixed) behind fftn(),
 * src/pyfourier/_fft/_fwd.py:18-196.
 * image (C, nimg, N...) with nimg = nframes (no basis) or ncoeff;
 * kspace (C, nframes, N...) in centred order; desc->nt must equal Nx (the
 * readout axis of b0_temporal); norm: PF_NORM_NONE (factor 1), _ORTHO
 * (1/sqrt(prod N)), _ADJOINT (1/prod N). */
int pf_cart_fft(pf_cart* plan, const pf_batch_desc* desc, int nframes, const void* image, void* kspace,
                int norm, void* workspace, size_t workspace_bytes, void* stream);

/* X_k = factor * sum_l conj(Ct_l) IFFTc(sum_t basis[t, k] conj(b_l(x)) mask_t kspace_t)
 * Replaces _fft_adj / _do_fft_adj / _basis_fwd, src/pyfourier/_fft/_fft.py:58-98,116-138
 * behind ifftn(), src/pyfourier/_fft/_adj.py:18-199.  norm: PF_NORM_NONE
 * (1/prod N), _ORTHO (1/sqrt(prod N)), _ADJOINT (1). */
int pf_cart_ifft(pf_cart* plan, const pf_batch_desc* desc, int nframes, const void* kspace, void* image,
                 int norm, void* workspace, size_t workspace_bytes, void* stream);

/* ---- GRAPPA operator gridding (GROG) ---------------------------------- */

/* Per-sample GROG bookkeeping.  Replaces the index / shift tables of
 * grog_interp, src/pyfourier/_preinterp/_grog.py:176-218.
 * coord: (npts, ndim) float32 grid units, last axis (x, y[, z]);
 * shape: host (ndim,) matrix (z, y, x); nsteps: odd quantisation levels.
 * lut:   (npts) int32: packed steps sx + sy*nsteps [+ sz*nsteps^2] with
 *        s_a = floor((nsteps-1) * (round(c_a) - c_a)) + nsteps/2, or -1 for a
 *        sample whose nearest grid point lies outside the matrix;
 * index: (npts, ndim) int32 round(c) + N/2, clamped into [0, N);
 * flat:  (npts) int64 sum_a index_a * prod_{b<a} N_b before clamping (the
 *        reference's key for counting repeated grid points). */
int pf_grog_index(const float* coord, int ndim, const int* shape, long long npts, int nsteps, int* lut, int* index,
                  long long* flat, void* stream);

/* Move every sample's coil vector onto its grid point:
 *   out[b, :, s] = in[b, :, s] @ T[0, s_x] @ T[1, s_y] [@ T[2, s_z]]
 * (zero where lut[s] = -1).  Replaces do_interpolation / _interp,
 * src/pyfourier/_preinterp/_grog.py:354-425, with the product operator of
 * :140-174 applied factor by factor.
 * in, out: (batch, ncoil, npts) complex64, not aliased; ncoil <= 64.
 * tables:  (ndim, nsteps, table_batch, ncoil, ncoil) complex64: per axis
 *          (x, y, z) and step the fractional operator power, row-vector
 *          convention (target = source @ G); table_batch is 1 (shared) or
 *          batch (one operator set per batch entry, e.g. per slice). */
int pf_grog_apply(const void* in, void* out, int batch, int ncoil, long long npts, int ndim, int nsteps,
                  const void* tables, int table_batch, const int* lut, void* stream);

/* ---- Voronoi density compensation ------------------------------------ */

/* Area of every 2-D Voronoi cell, 0 for open cells (points on the convex
 * hull).  Replaces the tessellation and cell areas of voronoi's _voronoi /
 * _voronoin / _compute_area, src/pyfourier/_dcomp/voronoi.py:220-258,277-343
 * (scipy.spatial.Voronoi + the polygon area), for distinct points.
 * pts:          (npoints, 2) float64, sorted by bucket of a uniform grid of
 *               nx * ny square buckets of side h with origin (x0, y0)
 *               (bucket = min(floor((y-y0)/h), ny-1)*nx + min(floor((x-x0)/h), nx-1));
 * bucket_start: (nx*ny + 1) int32 first point of each bucket, last = npoints;
 * area:         (npoints) float64, in the sorted order.
 * open:         (npoints) uint8 or NULL: 1 marks a point known to lie on the
 *               convex hull (vertex or edge), given area 0 without a scan.
 * list == NULL: every point, with a 64-vertex working polygon; a cell that
 *   needs more vertices gets area -1.  list != NULL: the nlist listed
 *   points again with work_vertices-vertex polygons (even, >= 4) in `work`
 *   (nlist * work_vertices * 40 bytes, 16-byte aligned); a cell that still
 *   does not fit sets *status to 1 (int32, cleared by the caller). */
int pf_voronoi_area_2d(const double* pts, const int* bucket_start, int nx, int ny, double x0, double y0, double h,
                       long long npoints, const unsigned char* open, const int* list, long long nlist, void* work,
                       int work_vertices, double* area, int* status, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PFHIP_H */
