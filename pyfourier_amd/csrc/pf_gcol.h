// Window geometry of the LDS-staged gathers (forward gather in pf_ops.hip,
// coordinate-derivative gather in pf_grad.hip) and their global -> LDS copy.
// HIP device code: include from .hip sources only.
#pragma once

#include "pf_internal.h"

namespace {

constexpr int GCOL_TB = PF_GCOL_MAXS;
constexpr int GCOL_OUT = 4;  // coil pairs per output store of the forward gather

template <int W>
struct GColShape {
    static constexpr int DZ = pf_gather_dz(W);
    static constexpr int NPL = DZ + W - 1;                          // staged z planes
    static constexpr int PX = PF_GCOL + W - 1;                       // staged points per y and x
    static constexpr int PP = PX * PX;
    static constexpr int NPTS = NPL * PP;
    static constexpr int NK = (NPTS + GCOL_TB - 1) / GCOL_TB;       // staged points per thread
};

// 16 bytes global -> LDS for each active lane (the destination is the wave's
// LDS base plus 16 bytes per lane)
__device__ __forceinline__ void gcol_copy16(const float4* src, float4* lds_wave_base) {
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_wave_base, 16, 0, 0);
}

}  // namespace
