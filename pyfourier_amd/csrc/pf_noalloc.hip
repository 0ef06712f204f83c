// Operator calls that allocate nothing: pf_nufft_noalloc.
//
// pf_nufft / pf_nufft_adj take a caller-allocated workspace (grid, tile-sorted
// samples, FFT intermediate, spread partial tiles).  Callers that cannot
// allocate device memory at call time (code under graph capture, tight
// serving loops, harnesses that forbid hipMalloc) use this entry point
// instead: the workspace is a fixed device array of this library, one copy per
// GPU, held until the process ends.
//
// The array's host-side name is not a device address.  Its address on each
// device is looked up once, with that device current, when the first plan is
// created there (pf_noalloc_prepare, called by every pf_plan_create*), kept
// per device, and only that pointer is ever passed to kernels.
#include <array>
#include <mutex>
#include <string>

#include "pf_internal.h"

static __device__ __align__(256) unsigned char scratch[PF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
std::array<void*, MAX_DEVICES> g_ptr{};  // device -> address of that device's scratch

void* scratch_of(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[dev];
}

}  // namespace

int pf_noalloc_prepare() {
    int dev = 0;
    PF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return pf_fail(PF_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[dev]) return PF_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return pf_fail(PF_ERR_HIP, std::string("scratch symbol address: ") + hipGetErrorString(e));
    g_ptr[dev] = ptr;
    return PF_OK;
}

extern "C" int pf_nufft_noalloc(pf_plan* p, const pf_batch_desc* d, int op, const void* in, void* out,
                                const float* weight, int norm, void* stream) {
    if (!p) return pf_fail(PF_ERR_ARG, "null plan");
    if (op != 0 && op != 1) return pf_fail(PF_ERR_ARG, "op must be 0 (forward) or 1 (adjoint)");
    if (!pf_fftk_supported(p))
        return pf_fail(PF_ERR_UNSUPPORTED,
                       "pf_nufft_noalloc needs grid sizes 2^a 3^b 5^c (other sizes run rocFFT, which allocates)");
    int dev = -1;
    PF_HIP(hipGetDevice(&dev));
    if (dev != p->device) return pf_fail(PF_ERR_ARG, "the plan lives on another device than the current one");
    size_t need = 0;
    int rc = pf_workspace_size(p, d, op, &need);
    if (rc != PF_OK) return rc;
    if (need > PF_NOALLOC_SCRATCH_BYTES)
        return pf_fail(PF_ERR_UNSUPPORTED, "pf_nufft_noalloc: the call needs " + std::to_string(need) +
                                               " bytes of workspace, more than its fixed " +
                                               std::to_string((size_t)PF_NOALLOC_SCRATCH_BYTES) +
                                               " (use pf_nufft / pf_nufft_adj with a workspace)");
    // looked up when the plan was created on this device; null only if that failed
    void* ws = scratch_of(dev);
    if (!ws) return pf_fail(PF_ERR_HIP, "the scratch address lookup failed on this device");
    // nothing in the workspace needs clearing: every kernel writes all of its
    // part (the spread writes every tile of the grid, empty ones as zeros)
    if (op == 0) return pf_nufft(p, d, in, out, weight, norm, ws, PF_NOALLOC_SCRATCH_BYTES, stream);
    return pf_nufft_adj(p, d, in, out, weight, norm, ws, PF_NOALLOC_SCRATCH_BYTES, stream);
}
