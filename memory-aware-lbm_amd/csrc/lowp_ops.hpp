// lowp_ops.hpp -- node operators on stored deviations g_i = f_i - w_i (the
// reduced-precision storage forms).  FP32 storage: the collision runs in FP64
// and each result is rounded to FP32 stochastically.  FP16 storage: the
// collision runs in FP32 and results are rounded to nearest.  Every fused
// multiply-add is written out (fma / fmaf) and the compiler forms no others
// (contraction off): which product a contracting compiler fuses depends on
// how often each product is used, i.e. on the code around the collision, and
// a kernel that needs one population of a collision would round differently
// from one that needs all nine.  Written out, every kernel computes the same
// bits.
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

template <int I> __device__ __forceinline__ double cdotu_d(double ux, double uy) {
    if constexpr (I == 0) return 0.;
    else if constexpr (I == 1) return ux;
    else if constexpr (I == 2) return uy;
    else if constexpr (I == 3) return -ux;
    else if constexpr (I == 4) return -uy;
    else if constexpr (I == 5) return ux + uy;
    else if constexpr (I == 6) return uy - ux;
    else if constexpr (I == 7) return -ux - uy;
    else return ux - uy;
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
#pragma clang fp contract(off)
    float north = g[2] + g[5] + g[6];
    float mid = g[0] + g[1] + g[3];
    float south = g[4] + g[7] + g[8];
    float drho = north + mid + south;
    float jx = (g[1] + g[5] + g[8]) - (g[3] + g[6] + g[7]);
    float jy = north - south;
    float shrink = drho / (1.f + drho);
    float ux = fmaf(-jx, shrink, jx), uy = fmaf(-jy, shrink, jy);
    float usq = fmaf(ux, ux, uy * uy);
    float u15 = -1.5f * usq;
    // moving populations first; the rest population's equilibrium closes the
    // mass balance exactly (sum geq = drho), so FP32 rounding of the
    // equilibrium cannot bias the density step after step
    float rest = drho;
    for_pops([&]<int I>() {
        if constexpr (I != 0) {
            float cu = cdotu_f<I>(ux, uy);
            float X = fmaf(4.5f * cu, cu, fmaf(3.f, cu, u15));
            // the weights 1/9 and 1/36 are not FP32 numbers: a one-term
            // product would scale every moving equilibrium by the same
            // (1 + 7.45e-9) and push that much extra momentum into the flow
            // at every collision, a drift that grows linearly in time; the
            // two-term weight leaves only unbiased rounding
            constexpr float w_hi = (float)wgt(I);
            constexpr float w_lo = (float)(wgt(I) - (double)w_hi);
            float Y = drho + fmaf(drho, X, X);
            float geq = fmaf(Y, w_hi, Y * w_lo);
            rest -= geq;
            g[I] = fmaf(omega, geq - g[I], g[I]);
        }
    });
    g[0] = fmaf(omega, rest - g[0], g[0]);
}

// ---- FP32 storage: FP64 collision, stochastic rounding --------------------------
//
// Round-to-nearest storage freezes the same rounding error into every node
// whose state repeats from step to step; in the large, nearly uniform regions
// of a big channel those errors are coherent and act as a steady forcing.
// On 16384 x 4096 after 20000 steps (Re 100) the velocity error against FP64
// is 3.0e-5 with FP32 collisions, 2.9e-5 with FP64 collisions rounded to
// nearest, 2.6e-5 with FP32 collisions rounded stochastically, and 5.2e-8
// with FP64 collisions (walls included) rounded stochastically.  The random
// bits are a hash of (node, collision index, population), so a run is
// deterministic and every kernel rounds a given node update identically; the
// node enters as its index in the reference's (lx+2) x (ly+2) arrays
// (Layout::key), not as a device address, so the bits do not depend on how
// the device lays the populations out.

__device__ __forceinline__ uint32_t sr_mix(uint32_t a) {
    a ^= a >> 16;
    a *= 0x7feb352du;
    a ^= a >> 15;
    a *= 0x846ca68bu;
    a ^= a >> 16;
    return a;
}

// 64 random bits per node and collision (two mixing rounds of the node index
// and the collision index); each population takes its own 7 of them
struct SrSeed { uint32_t lo, hi; };
__device__ __forceinline__ SrSeed sr_seed(int64_t key, uint32_t step) {
    const uint32_t a = sr_mix((uint32_t)key * 0x9e3779b1u + step * 0x85ebca77u);
    return {a, sr_mix(a ^ 0x632be5abu)};
}

// v rounded to one of its two FP32 neighbours with probability proportional
// to closeness: a random offset is added below FP32's last mantissa bit and
// the sum is truncated toward zero (IEEE is sign-magnitude, so the same works
// for both signs).  The offset is (2k + 1) * 2^21 for the population's 7
// random bits k: the rounding probability is resolved to 1/128 and, with the
// half step, unbiased on average.  RZ: the kernel runs with FP32 results
// rounding toward zero (fp32_round_toward_zero), so the conversion truncates
// by itself; otherwise a round-to-nearest conversion is corrected.  Both give
// the same bits.
template <int I, bool RZ> __device__ __forceinline__ float sr_f32(double v, SrSeed h) {
    uint32_t k;
    if constexpr (7 * I + 7 <= 32) k = (h.lo >> (7 * I)) & 0x7fu;
    else if constexpr (7 * I >= 32) k = (h.hi >> (7 * I - 32)) & 0x7fu;
    else k = __builtin_amdgcn_alignbit(h.hi, h.lo, 7 * I) & 0x7fu;
    const uint64_t off = (uint64_t)(2u * k + 1u) << 21;
    const double vb = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) + off);
    float f = (float)vb;
    if constexpr (!RZ)
        if (fabs((double)f) > fabs(vb)) f = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, f) - 1u);
    return f;
}

// Switch this wave's FP32 rounding to toward-zero (MODE.FP_ROUND bits 1:0;
// FP64 and FP16 keep rounding to nearest).  Kernels that call it do no FP32
// arithmetic afterwards other than the conversions of sr_f32<true>, and no
// integer division (whose expansion uses FP32 reciprocals).
__device__ __forceinline__ void fp32_round_toward_zero() {
    __builtin_amdgcn_s_setreg(0x1801 /* hwreg(HW_REG_MODE, 0, 4) */, 3);
}

// BGK on FP32-stored deviations, computed in FP64 (same deviation form as
// bgk_dev; the mass-closing rest population is kept), rounded stochastically.
template <bool RZ>
__device__ __forceinline__ void bgk_f32s(float g[9], double omega, SrSeed h) {
#pragma clang fp contract(off)
    double d[9];
    for_pops([&]<int I>() { d[I] = g[I]; });
    const double north = d[2] + d[5] + d[6], mid = d[0] + d[1] + d[3], south = d[4] + d[7] + d[8];
    const double drho = north + mid + south;
    // 1 / (1 + drho): hardware reciprocal and one Newton step (a rounding or
    // two off the correctly rounded quotient; the same bits in every kernel)
    const double a = 1.0 + drho;
    double inv = __builtin_amdgcn_rcp(a);
    inv = fma(inv, fma(-a, inv, 1.0), inv);
    const double ux = ((d[1] + d[5] + d[8]) - (d[3] + d[6] + d[7])) * inv;
    const double uy = (north - south) * inv;
    const double usq = fma(ux, ux, uy * uy);
    const double u15 = -1.5 * usq;
    double rest = drho;
    for_pops([&]<int I>() {
        if constexpr (I != 0) {
            const double cu = cdotu_d<I>(ux, uy);
            const double X = fma(4.5 * cu, cu, fma(3.0, cu, u15));
            const double geq = wgt(I) * (drho + fma(drho, X, X));
            rest -= geq;
            g[I] = sr_f32<I, RZ>(fma(omega, geq - d[I], d[I]), h);
        }
    });
    g[0] = sr_f32<0, RZ>(fma(omega, rest - d[0], d[0]), h);
}

// a wall node's FP64 result back to the storage format
template <typename S, int I, bool RZ> __device__ __forceinline__ float to_dev(double f, SrSeed h) {
    if constexpr (sizeof(S) == 4) return sr_f32<I, RZ>(f - wgt(I), h);
    else return (float)(f - wgt(I));
}

// the bulk collision of either reduced format
template <typename S, bool RZ>
__device__ __forceinline__ void bgk_lp(float g[9], double omega, SrSeed h) {
    if constexpr (sizeof(S) == 4) bgk_f32s<RZ>(g, omega, h);
    else bgk_dev(g, (float)omega);
}

__device__ __forceinline__ void bounce_dev(float g[9]) {
    float t;
    t = g[1]; g[1] = g[3]; g[3] = t;
    t = g[2]; g[2] = g[4]; g[4] = t;
    t = g[5]; g[5] = g[7]; g[7] = t;
    t = g[6]; g[6] = g[8]; g[8] = t;
}

}  // namespace lbm
