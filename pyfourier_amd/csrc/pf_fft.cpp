// Cartesian FFT stage on the plan's oversampled grid with rocFFT.
//
// The grid slab of one (frame grid, coil block) is (G..., R*Cv) with the
// R*Cv batch values innermost, so each transform is an ndim-D c2c FFT with
// element stride R*Cv and batch distance 1 (rocFFT advanced data layout).
// Transforms are unnormalised in both directions; the operator folds the
// reference's norm=None / "ortho" factors and 1/scale into one multiply
// (src/pyfourier/_subroutines/_fft.py:72-79,163-170).  Shift pairs are folded
// into the index maps of the pad/crop and gather/spread kernels.
#include <rocfft/rocfft.h>

#include <map>
#include <new>
#include <mutex>
#include <tuple>

#include "pf_internal.h"

struct pf_fft_entry {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
};

struct pf_fft_cache {
    std::map<std::tuple<int, int>, pf_fft_entry> entries;  // (batch RC, inverse)
    std::mutex mu;
};

static std::once_flag g_rocfft_once;
static rocfft_status g_rocfft_setup = rocfft_status_success;

void pf_fft_destroy(pf_fft_cache* c) {
    if (!c) return;
    for (auto& kv : c->entries) {
        if (kv.second.info) rocfft_execution_info_destroy(kv.second.info);
        if (kv.second.plan) rocfft_plan_destroy(kv.second.plan);
        if (kv.second.work) (void)hipFree(kv.second.work);
    }
    delete c;
}

int pf_fft_cache_init(pf_plan* p) {
    std::call_once(g_rocfft_once, [] { g_rocfft_setup = rocfft_setup(); });
    if (g_rocfft_setup != rocfft_status_success) return pf_fail(PF_ERR_FFT, "rocfft_setup failed");
    if (!p->fft) p->fft = new (std::nothrow) pf_fft_cache();
    if (!p->fft) return pf_fail(PF_ERR_ALLOC, "host allocation failed");
    return PF_OK;
}

static int get_entry(pf_plan* p, int RC, int inverse, pf_fft_entry** out) {
    // the cache exists from plan creation, so concurrent first calls only
    // meet under its lock
    if (!p->fft) {
        int rc = pf_fft_cache_init(p);
        if (rc != PF_OK) return rc;
    }
    std::lock_guard<std::mutex> lk(p->fft->mu);
    auto key = std::make_tuple(RC, inverse);
    auto it = p->fft->entries.find(key);
    if (it != p->fft->entries.end()) {
        *out = &it->second;
        return PF_OK;
    }
    const int d = p->ndim;
    size_t lengths[3], strides[3];
    // rocFFT lists dimensions fastest first: x, y, z
    size_t st = (size_t)RC;
    for (int i = 0; i < d; ++i) {
        lengths[i] = (size_t)p->os[2 - i];
        strides[i] = st;
        st *= lengths[i];
    }
    rocfft_plan_description desc = nullptr;
    if (rocfft_plan_description_create(&desc) != rocfft_status_success)
        return pf_fail(PF_ERR_FFT, "rocfft_plan_description_create failed");
    rocfft_status s = rocfft_plan_description_set_data_layout(
        desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr, nullptr,
        (size_t)d, strides, 1, (size_t)d, strides, 1);
    if (s != rocfft_status_success) {
        rocfft_plan_description_destroy(desc);
        return pf_fail(PF_ERR_FFT, "rocfft_plan_description_set_data_layout failed");
    }
    pf_fft_entry e;
    s = rocfft_plan_create(&e.plan, rocfft_placement_inplace,
                           inverse ? rocfft_transform_type_complex_inverse : rocfft_transform_type_complex_forward,
                           rocfft_precision_single, (size_t)d, lengths, (size_t)RC, desc);
    rocfft_plan_description_destroy(desc);
    if (s != rocfft_status_success) return pf_fail(PF_ERR_FFT, "rocfft_plan_create failed");
    rocfft_plan_get_work_buffer_size(e.plan, &e.work_bytes);
    if (rocfft_execution_info_create(&e.info) != rocfft_status_success) {
        rocfft_plan_destroy(e.plan);
        return pf_fail(PF_ERR_FFT, "rocfft_execution_info_create failed");
    }
    if (e.work_bytes) {
        if (hipMalloc(&e.work, e.work_bytes) != hipSuccess) {
            rocfft_execution_info_destroy(e.info);
            rocfft_plan_destroy(e.plan);
            return pf_fail(PF_ERR_ALLOC, "rocFFT work buffer allocation failed");
        }
        rocfft_execution_info_set_work_buffer(e.info, e.work, e.work_bytes);
    }
    auto res = p->fft->entries.emplace(key, e);
    *out = &res.first->second;
    return PF_OK;
}

int pf_fft_exec(pf_plan* p, const pf_batch& b, float2* grid, int inverse, hipStream_t stream) {
    const int RC = b.R * b.Cv;
    pf_fft_entry* e = nullptr;
    int rc = get_entry(p, RC, inverse, &e);
    if (rc != PF_OK) return rc;
    if (rocfft_execution_info_set_stream(e->info, stream) != rocfft_status_success)
        return pf_fail(PF_ERR_FFT, "rocfft_execution_info_set_stream failed");
    const int64_t slab = pf_gtot(p) * RC;
    for (int s = 0; s < b.nfg * b.Cb; ++s) {
        void* buf[1] = {grid + (int64_t)s * slab};
        if (rocfft_execute(e->plan, buf, nullptr, e->info) != rocfft_status_success)
            return pf_fail(PF_ERR_FFT, "rocfft_execute failed");
    }
    return PF_OK;
}
