// collide_f64.hpp -- one node's FP64 collision in the cylinder channel,
// shared by the single-step and two-step kernels.
#pragma once

#include "d2q9_fp64.hpp"
#include "lbm_common.hpp"

namespace lbm {

__device__ __forceinline__ bool outside(const Layout &L, int x, int y) {
    return x < 1 || x > L.lx || y < 1 || y > L.ly;
}

template <int S>
__device__ __forceinline__ void wall_side(const Problem &p, int x, int y, double f[9],
                                          double right_p0) {
    const SideBC &b = p.side[S];
    const int pos = (S == UPPER || S == LOWER) ? x : y;
    double p0 = b.p0[pos], p1 = b.p1[pos];
    if constexpr (S == RIGHT) {
        if (p.zgb) p0 = right_p0;
    }
    wall_node<S>(f, b.kind, b.family, p0, p1, p.omega);
}

// The reference's per-node dispatch (lb.c:125-128) resolved by class:
// side s >= 0 -> that side's straight-wall operator (its p0 replaced by
// right_p0 on a zero-gradient outlet), disc -> bounce-back, else BGK.
__device__ __forceinline__ void collide_class(const Problem &p, int s, int x, int y,
                                              double f[9], double right_p0) {
    switch (s) {
    case UPPER: wall_side<UPPER>(p, x, y, f, right_p0); break;
    case LOWER: wall_side<LOWER>(p, x, y, f, right_p0); break;
    case LEFT: wall_side<LEFT>(p, x, y, f, right_p0); break;
    case RIGHT: wall_side<RIGHT>(p, x, y, f, right_p0); break;
    default:
        if (p.r2 >= 0 && in_disc(p, x, y)) bounce_back(f);
        else bgk(f, p.omega);
    }
}

// zero-gradient outlet density from two post-stream densities
// (updateZeroGradientBoundary, unsteady.c:211)
__device__ __forceinline__ double zgb_rho(double rho1, double rho2) {
    FP64_EXACT;
    return 4. / 3. * rho1 - 1. / 3. * rho2;
}

}  // namespace lbm
