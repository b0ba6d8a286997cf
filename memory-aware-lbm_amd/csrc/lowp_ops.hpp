// lowp_ops.hpp -- FP32 node operators on stored deviations g_i = f_i - w_i
// (the reduced-precision storage forms).  Fused multiply-add is allowed here
// whatever the translation unit's -ffp-contract, so every kernel that
// includes this header computes the same bits.
#pragma once

#include "d2q9_fp64.hpp"

namespace lbm {

// Narrow an FP32 result to the storage type.  For FP16 the value is pinned in
// a register first: otherwise the compiler may fold the last FP32 fma and the
// narrowing into one v_fma_mixlo_f16 (a single rounding instead of FP32 then
// FP16), in some kernels and not in others, and the same node would get
// different bits depending on the code around it.
template <typename S> __device__ __forceinline__ S narrow(float v) {
    if constexpr (sizeof(S) == 2) __asm__ volatile("" : "+v"(v));
    return (S)v;
}

template <int I> __device__ __forceinline__ float cdotu_f(float ux, float uy) {
    if constexpr (I == 0) return 0.f;
    else if constexpr (I == 1) return ux;
    else if constexpr (I == 2) return uy;
    else if constexpr (I == 3) return -ux;
    else if constexpr (I == 4) return -uy;
    else if constexpr (I == 5) return ux + uy;
    else if constexpr (I == 6) return uy - ux;
    else if constexpr (I == 7) return -ux - uy;
    else return ux - uy;
}

// BGK on deviations, FP32.  rho = 1 + drho is never formed in FP32 (it would
// drop the low bits of drho every step): u = j / (1 + drho) is evaluated as
// j - j drho / (1 + drho) and rho X as X + drho X, so the rounding of 1 + drho
// only touches terms that are already O(drho) small.
__device__ __forceinline__ void bgk_dev(float g[9], float omega) {
#pragma clang fp contract(fast)
    float north = g[2] + g[5] + g[6];
    float mid = g[0] + g[1] + g[3];
    float south = g[4] + g[7] + g[8];
    float drho = north + mid + south;
    float jx = (g[1] + g[5] + g[8]) - (g[3] + g[6] + g[7]);
    float jy = north - south;
    float shrink = drho / (1.f + drho);
    float ux = jx - jx * shrink, uy = jy - jy * shrink;
    float usq = ux * ux + uy * uy;
    // moving populations first; the rest population's equilibrium closes the
    // mass balance exactly (sum geq = drho), so FP32 rounding of the
    // equilibrium cannot bias the density step after step
    float rest = drho;
    for_pops([&]<int I>() {
        if constexpr (I != 0) {
#pragma clang fp contract(fast)
            float cu = cdotu_f<I>(ux, uy);
            float X = 3.f * cu + 4.5f * cu * cu - 1.5f * usq;
            float geq = (float)wgt(I) * (drho + (X + drho * X));
            rest -= geq;
            g[I] += omega * (geq - g[I]);
        }
    });
    g[0] += omega * (rest - g[0]);
}

__device__ __forceinline__ void bounce_dev(float g[9]) {
    float t;
    t = g[1]; g[1] = g[3]; g[3] = t;
    t = g[2]; g[2] = g[4]; g[4] = t;
    t = g[5]; g[5] = g[7]; g[7] = t;
    t = g[6]; g[6] = g[8]; g[8] = t;
}

}  // namespace lbm
