// node_step.hpp -- one node's full step (pull or local read, collision with
// the stale-slot and outlet bookkeeping, store), for FP64 populations and for
// the FP32/FP16 deviation storage.  Shared by the single-lattice kernels and
// the batched persistent kernel, so both compute the same bits.
#pragma once

#include "collide_f64.hpp"
#include "lowp_ops.hpp"

namespace lbm {

// Post-stream density of interior node (x, y): gathered from a collided
// buffer, or read in place from a streamed one.
template <bool PULL>
__device__ __forceinline__ double density_at(const Layout &L, const double *__restrict__ s,
                                             int x, int y) {
    double g[9];
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() {
        g[I] = PULL ? s[I * L.plane + n - cx(I) * L.pitch - cy(I)] : s[I * L.plane + n];
    });
    return rho_of(g);
}

// Collision of one perimeter node, with the stale-slot bookkeeping and the
// zero-gradient outlet density.  Kept out of line: perimeter nodes are a
// thin minority and the bulk path should stay small.
template <bool PULL>
__device__ __forceinline__ void perimeter_node(const Problem &p, int s, int x, int y,
                                            double f[9], const double *__restrict__ src,
                                            double *__restrict__ dst) {
    const Layout &L = p.lay;
    int64_t n = L.node(x, y);
    // Populations entering this node from outside the domain must be present
    // again at the next step (the reference reads them at the four corners).
    // Collided state: keep them in dst's ghost ring = this node's value of
    // the step before (src(x)_i).  Streamed state: that value sits in dst(x)
    // itself and is read before dst(x) is overwritten below.
    for_pops([&]<int I>() {
        if constexpr (I != 0) {
            int gx = x - cx(I), gy = y - cy(I);
            if (outside(L, gx, gy)) {
                double prev = PULL ? src[I * L.plane + n] : dst[I * L.plane + n];
                dst[I * L.plane + L.node(gx, gy)] = prev;
            }
        }
    });
    double right_p0 = 0.;
    if (s == RIGHT && p.zgb) {
        // updateZeroGradientBoundary (unsteady.c:205-214) on the post-stream
        // state of this same step.
        right_p0 = zgb_rho(density_at<PULL>(L, src, L.lx - 1, y),
                           density_at<PULL>(L, src, L.lx - 2, y));
    }
    if (s == RIGHT) p.outlet_rho[y] = p.zgb ? right_p0 : p.side[RIGHT].p0[y];
    collide_class(p, s, x, y, f, right_p0);
}


// The collision of node (x, y) on its post-stream populations f (pulled from
// a collided src, or read in place from a streamed one), with the perimeter
// bookkeeping (which writes dst's ghost ring and, on a streamed state, reads
// dst(x, y) before the caller overwrites it).
template <bool PULL>
__device__ __forceinline__ void collide_node_f64(const Problem &p, int x, int y, double f[9],
                                                 const double *__restrict__ src,
                                                 double *__restrict__ dst) {
    int s = side_of(p, x, y);
    if (s >= 0) perimeter_node<PULL>(p, s, x, y, f, src, dst);
    else if (p.r2 >= 0 && in_disc(p, x, y)) bounce_back(f);
    else bgk(f, p.omega);
}

template <bool PULL>
__device__ __forceinline__ void node_step_f64(const Problem &p, int x, int y,
                                              const double *__restrict__ src,
                                              double *__restrict__ dst) {
    const Layout &L = p.lay;
    int64_t n = L.node(x, y);
    double f[9];
    for_pops([&]<int I>() {
        f[I] = PULL ? src[I * L.plane + n - cx(I) * L.pitch - cy(I)] : src[I * L.plane + n];
    });
    collide_node_f64<PULL>(p, x, y, f, src, dst);
    for_pops([&]<int I>() { __builtin_nontemporal_store(f[I], &dst[I * L.plane + n]); });
}

// ---- reduced-precision storage ------------------------------------------------

template <typename S> __device__ __forceinline__ float ld(const S *p) { return (float)*p; }
template <typename S> __device__ __forceinline__ void st_nt(S v, S *p) {
    __builtin_nontemporal_store(v, p);
}

template <typename S, bool PULL>
__device__ __forceinline__ double density_lp(const Layout &L, const S *__restrict__ s, int x,
                                             int y) {
    double f[9];
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() {
        float g = PULL ? ld(&s[I * L.plane + n - cx(I) * L.pitch - cy(I)]) : ld(&s[I * L.plane + n]);
        f[I] = wgt(I) + (double)g;
    });
    return rho_of(f);
}

template <typename S, bool PULL, bool RZ>
__device__ __forceinline__ void perimeter_lp(const Problem &p, int s, int x, int y, float g[9],
                                          const S *__restrict__ src, S *__restrict__ dst,
                                          SrSeed h) {
    const Layout &L = p.lay;
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() {
        if constexpr (I != 0) {
            int gx = x - cx(I), gy = y - cy(I);
            if (outside(L, gx, gy)) dst[I * L.plane + L.node(gx, gy)] = PULL ? src[I * L.plane + n]
                                                                             : dst[I * L.plane + n];
        }
    });
    double f[9];
    for_pops([&]<int I>() { f[I] = wgt(I) + (double)g[I]; });
    double rp0 = 0.;
    if (s == RIGHT && p.zgb)
        rp0 = zgb_rho(density_lp<S, PULL>(L, src, L.lx - 1, y), density_lp<S, PULL>(L, src, L.lx - 2, y));
    if (s == RIGHT) p.outlet_rho[y] = p.zgb ? rp0 : p.side[RIGHT].p0[y];
    collide_class(p, s, x, y, f, rp0);
    for_pops([&]<int I>() { g[I] = to_dev<S, I, RZ>(f[I], h); });
}

// The collision of node (x, y) on its post-stream deviations g, reduced
// storage (perimeter bookkeeping as collide_node_f64); `step` is the index of
// this collision (it keys FP32's stochastic rounding); RZ as in sr_f32.
template <typename S, bool PULL, bool RZ>
__device__ __forceinline__ void collide_node_lp(const Problem &p, int x, int y, float g[9],
                                                const S *__restrict__ src, S *__restrict__ dst,
                                                uint32_t step) {
    const SrSeed h = sizeof(S) == 4 ? sr_seed(p.lay.key(x, y), step) : SrSeed{0u, 0u};
    int s = side_of(p, x, y);
    if (s >= 0) perimeter_lp<S, PULL, RZ>(p, s, x, y, g, src, dst, h);
    else if (p.r2 >= 0 && in_disc(p, x, y)) bounce_dev(g);
    else bgk_lp<S, RZ>(g, p.omega, h);
}

// One collision of node (x, y) with reduced storage; `step` is the index of
// this collision (it keys FP32's stochastic rounding); RZ as in sr_f32.
template <typename S, bool PULL, bool RZ>
__device__ __forceinline__ void node_step_lp(const Problem &p, int x, int y,
                                             const S *__restrict__ src, S *__restrict__ dst,
                                             uint32_t step) {
    const Layout &L = p.lay;
    int64_t n = L.node(x, y);
    float g[9];
    for_pops([&]<int I>() {
        g[I] = PULL ? ld(&src[I * L.plane + n - cx(I) * L.pitch - cy(I)]) : ld(&src[I * L.plane + n]);
    });
    collide_node_lp<S, PULL, RZ>(p, x, y, g, src, dst, step);
    for_pops([&]<int I>() { st_nt(narrow<S>(g[I]), &dst[I * L.plane + n]); });
}

}  // namespace lbm
