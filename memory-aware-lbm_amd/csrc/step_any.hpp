// step_any.hpp -- one node's collision and the densities around it for
// every storage format, on compute-type values: FP64 populations for FP64
// storage, FP32 deviations from the weights for FP32 / FP16 storage.  Used by
// the multi-step kernels (kernels_wave.hip, block_tile.hpp), whose
// intermediate steps never reach memory: each is rounded to the storage
// format exactly as a store would round it, so k steps in one launch equal k
// single steps bit for bit.
#pragma once

#include "fields.hpp"
#include "node_step.hpp"

namespace lbm {

// Integer division expands into FP32 reciprocal arithmetic, which must not
// run under toward-zero FP32 rounding (fp32_round_toward_zero): an index
// computed by division is passed through here before the rounding mode
// changes, and the side-effect order keeps the division ahead of the switch.
__device__ __forceinline__ void pin_index(int &v) { __asm__ volatile("" : "+v"(v)); }

template <typename S> struct Calc { using T = float; };
template <> struct Calc<double> { using T = double; };

template <typename S> __device__ __forceinline__ typename Calc<S>::T as_calc(S v) {
    return (typename Calc<S>::T)v;
}

// round a computed value to what storage would hold, back in compute type
template <typename S> __device__ __forceinline__ typename Calc<S>::T stored(typename Calc<S>::T v) {
    if constexpr (sizeof(S) == 2) return (float)narrow<S>(v);
    else return v;
}

// density of compute-type populations, summed in FP64 the way the
// single-step kernels sum what they load
template <typename S> __device__ __forceinline__ double rho_of_any(const typename Calc<S>::T h[9]) {
    if constexpr (sizeof(S) == 8) {
        return rho_of(h);
    } else {
        double f[9];
        for_pops([&]<int I>() { f[I] = wgt(I) + (double)h[I]; });
        return rho_of(f);
    }
}

// post-stream density of interior node (x, y) gathered from P
template <typename S>
__device__ __forceinline__ double pulled_rho(const Layout &L, const S *__restrict__ P, int x, int y) {
    double f[9];
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() {
        f[I] = pop_f64<S, I>(&P[I * L.plane + n - cx(I) * L.pitch - cy(I)]);
    });
    return rho_of(f);
}

// One node's collision on compute-type populations h, side s (-1: interior);
// `step` is the collision's index (it keys FP32's stochastic rounding, so the
// first step held in registers rounds exactly as a stored single step would).
template <typename S>
__device__ __forceinline__ void collide_any(const Problem &p, int s, int x, int y,
                                            typename Calc<S>::T h[9], double right_p0,
                                            uint32_t step) {
    // kernels that call this run FP32 storage with toward-zero FP32 rounding
    // (fp32_round_toward_zero)
    if constexpr (sizeof(S) == 8) {
        collide_class(p, s, x, y, h, right_p0);
    } else {
        const SrSeed hs = sizeof(S) == 4 ? sr_seed(p.lay.key(x, y), step) : SrSeed{0u, 0u};
        if (s < 0) {
            if (p.r2 >= 0 && in_disc(p, x, y)) bounce_dev(h);
            else bgk_lp<S, true>(h, p.omega, hs);
            return;
        }
        double f[9];
        for_pops([&]<int I>() { f[I] = wgt(I) + (double)h[I]; });
        collide_class(p, s, x, y, f, right_p0);
        for_pops([&]<int I>() { h[I] = to_dev<S, I, true>(f[I], hs); });
    }
}

// Some node of columns [xa, xb] x rows [ya, yb] lies in the disc (exact:
// the box's point nearest the centre is tested).
__host__ __device__ inline bool box_meets_disc(const Problem &p, int xa, int xb, int ya, int yb) {
    if (p.r2 < 0) return false;
    const int64_t dx = p.ox < xa ? xa - p.ox : (p.ox > xb ? p.ox - xb : 0);
    const int64_t dy = p.oy < ya ? ya - p.oy : (p.oy > yb ? p.oy - yb : 0);
    return dx * dx + dy * dy <= p.r2;
}

// collide_any of an interior node (s = -1); DISC = false when the caller has
// established that the node lies outside the disc, which leaves the bulk
// collision alone, with no per-lane branch.
template <typename S, bool DISC>
__device__ __forceinline__ void collide_bulk(const Problem &p, int x, int y,
                                             typename Calc<S>::T h[9], uint32_t step) {
    if constexpr (DISC) {
        collide_any<S>(p, -1, x, y, h, 0., step);
    } else if constexpr (sizeof(S) == 8) {
        bgk(h, p.omega);
    } else {
        const SrSeed hs = sizeof(S) == 4 ? sr_seed(p.lay.key(x, y), step) : SrSeed{0u, 0u};
        bgk_lp<S, true>(h, p.omega, hs);
    }
}

// a computed value as the storage format holds it
template <typename S>
__device__ __forceinline__ S to_storage(typename Calc<S>::T v) {
    if constexpr (sizeof(S) == 8) return v;
    else return narrow<S>(v);
}

}  // namespace lbm
