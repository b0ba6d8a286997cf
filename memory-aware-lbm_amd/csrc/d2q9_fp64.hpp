// d2q9_fp64.hpp -- double-precision D2Q9 node operators for gfx950.
//
// These are the per-node numerics of the collide-stream path: moments,
// equilibrium, BGK relaxation, full-way bounce-back and the straight-wall
// Zou/He and regularized operators.  Every expression keeps the evaluation
// order of the reference C code it replaces (cited per function) and this
// header is only compiled with -ffp-contract=off, so each double operation
// rounds exactly once, as in the reference's SSE2 build.  That is what makes
// the device results bit-identical to the reference.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

// Every FP64 operator below rounds each operation exactly once, whatever
// -ffp-contract the including translation unit uses.
#define FP64_EXACT _Pragma("clang fp contract(off)")

// D2Q9 lattice (lb.c:48-55): 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE.
__host__ __device__ constexpr int cx(int i) {
    return (i == 1 || i == 5 || i == 8) ? 1 : (i == 3 || i == 6 || i == 7) ? -1 : 0;
}
__host__ __device__ constexpr int cy(int i) {
    return (i == 2 || i == 5 || i == 6) ? 1 : (i == 4 || i == 7 || i == 8) ? -1 : 0;
}
__host__ __device__ constexpr double wgt(int i) {
    return i == 0 ? 4. / 9. : (i <= 4 ? 1. / 9. : 1. / 36.);
}
__host__ __device__ constexpr int opp(int i) {  // boundaries.c:32
    return i == 0 ? 0 : (i <= 4 ? ((i + 1) & 3) + 1 : ((i - 3) & 3) + 5);
}

enum Side : int { UPPER = 0, LOWER = 1, LEFT = 2, RIGHT = 3 };

// Moments (lb.c:134-141): density summed by lines of constant c_y; the
// velocity by true IEEE division.
__device__ __forceinline__ double rho_of(const double f[9]) {
    FP64_EXACT;
    double north = f[2] + f[5] + f[6];
    double mid = f[0] + f[1] + f[3];
    double south = f[4] + f[7] + f[8];
    return north + mid + south;
}

// a / d and b / d, each the correctly rounded IEEE quotient: the hardware
// division sequence (scale, reciprocal and two Newton steps on the
// denominator, then per quotient a product, its remainder, the fused
// correction and the special-case fixup), with the denominator's part done
// once for both quotients.  The denominator's own scale step takes d as the
// numerator too, which leaves a normal d unscaled, the value the sequence of
// a / d uses whenever a is zero or at least 2^-969 in magnitude; populations
// of a lattice give numerators that are zero or far larger than that, so
// both quotients have the bits of a / d and b / d.
__device__ __forceinline__ void div_same_den(double a, double b, double d, double &qa,
                                             double &qb) {
    FP64_EXACT;
    bool unused;
    const double ds = __builtin_amdgcn_div_scale(d, d, false, &unused);
    const double r0 = __builtin_amdgcn_rcp(ds);
    const double r1 = __builtin_fma(r0, __builtin_fma(-ds, r0, 1.0), r0);
    const double r2 = __builtin_fma(r1, __builtin_fma(-ds, r1, 1.0), r1);
    auto quot = [&](double n) {
        bool sc;
        const double ns = __builtin_amdgcn_div_scale(n, d, true, &sc);
        const double m = ns * r2;
        const double rem = __builtin_fma(-ds, m, ns);
        return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(rem, r2, m, sc), d, n);
    };
    qa = quot(a);
    qb = quot(b);
}

__device__ __forceinline__ void macros(const double f[9], double &rho, double &ux,
                                       double &uy) {
    FP64_EXACT;
    double north = f[2] + f[5] + f[6];
    double mid = f[0] + f[1] + f[3];
    double south = f[4] + f[7] + f[8];
    rho = north + mid + south;
    // ux = (f1 + f5 + f8 - (f3 + f6 + f7)) / rho, uy = (north - south) / rho
    div_same_den(f[1] + f[5] + f[8] - (f[3] + f[6] + f[7]), north - south, rho, ux, uy);
}

// Equilibrium (lb.c:144-151).  c.u is formed from integer components; the
// products by +-1 and 0 are exact, so the specialised sums below round the
// same way as cx*ux + cy*uy.
template <int I>
__device__ __forceinline__ double cdotu(double ux, double uy) {
    FP64_EXACT;
    if constexpr (I == 0) return 0.;
    else if constexpr (I == 1) return ux;
    else if constexpr (I == 2) return uy;
    else if constexpr (I == 3) return -ux;
    else if constexpr (I == 4) return -uy;
    else if constexpr (I == 5) return ux + uy;
    else if constexpr (I == 6) return -ux + uy;
    else if constexpr (I == 7) return -ux + -uy;
    else return ux + -uy;
}

template <int I>
__device__ __forceinline__ double feq(double rho, double ux, double uy, double usqr) {
    FP64_EXACT;
    double cu = cdotu<I>(ux, uy);
    return rho * wgt(I) * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usqr);
}

template <int... Is> struct seq {};
template <typename F, int... Is>
__device__ __forceinline__ void for_pops_impl(F &&fn, seq<Is...>) {
    FP64_EXACT;
    (fn.template operator()<Is>(), ...);
}
template <typename F> __device__ __forceinline__ void for_pops(F &&fn) {
    FP64_EXACT;
    for_pops_impl(fn, seq<0, 1, 2, 3, 4, 5, 6, 7, 8>{});
}

// BGK relaxation (lb.c:154-166): f_i = f_i*(1-omega), then += omega*feq_i.
__device__ __forceinline__ void bgk(double f[9], double omega) {
    FP64_EXACT;
    double rho, ux, uy;
    macros(f, rho, ux, uy);
    double usqr = ux * ux + uy * uy;
    double keep = 1 - omega;
    for_pops([&]<int I>() {
        FP64_EXACT;
        f[I] *= keep;
        f[I] += omega * feq<I>(rho, ux, uy, usqr);
    });
}

// Full-way bounce-back (boundaries.c:48-59).
__device__ __forceinline__ void bounce_back(double f[9]) {
    FP64_EXACT;
    double t;
    t = f[1]; f[1] = f[3]; f[3] = t;
    t = f[2]; f[2] = f[4]; f[4] = t;
    t = f[5]; f[5] = f[7]; f[7] = t;
    t = f[6]; f[6] = f[8]; f[8] = t;
}

// The wall operators below take the side as a template parameter: with a
// run-time side the compiler turns the per-side population indices into a
// lookup table and the population array into scratch memory.

// Known-population sum at a wall node (boundaries.c:66-126 bracket it as
// tangential triple plus twice the outgoing triple).
template <int S> __device__ __forceinline__ double wall_sum(const double f[9]) {
    FP64_EXACT;
    if constexpr (S == UPPER) return f[0] + f[3] + f[1] + 2 * (f[6] + f[2] + f[5]);
    else if constexpr (S == LOWER) return f[0] + f[3] + f[1] + 2 * (f[8] + f[4] + f[7]);
    else if constexpr (S == RIGHT) return f[0] + f[2] + f[4] + 2 * (f[1] + f[5] + f[8]);
    else return f[0] + f[2] + f[4] + 2 * (f[3] + f[6] + f[7]);
}

// Density from a prescribed normal velocity (upperRho, lowerRho, rightRho,
// leftRho: boundaries.c:66-118): reciprocal first, then the product.
template <int S> __device__ __forceinline__ double wall_rho(const double f[9], double un) {
    FP64_EXACT;
    double den = (S == UPPER || S == RIGHT) ? 1. + un : 1. - un;
    return 1. / den * wall_sum<S>(f);
}

// Normal velocity from a prescribed density (upperU, lowerU, rightU, leftU:
// boundaries.c:74-126).
template <int S> __device__ __forceinline__ double wall_un(const double f[9], double rho) {
    FP64_EXACT;
    double q = 1. / rho * wall_sum<S>(f);
    return (S == UPPER || S == RIGHT) ? -1. + q : 1. - q;
}

// Zou/He completion of the three unknown populations (boundaries.c:131-169).
template <int S>
__device__ __forceinline__ void zh_complete(double f[9], double ux, double uy, double rho) {
    FP64_EXACT;
    if constexpr (S == UPPER) {
        f[7] = f[5] + 0.5 * (f[1] - f[3]) - rho * uy / 6 - rho * ux / 2.;
        f[8] = f[6] + 0.5 * (f[3] - f[1]) - rho * uy / 6 + rho * ux / 2.;
        f[4] = f[2] - 2. / 3. * rho * uy;
    } else if constexpr (S == LOWER) {
        f[6] = f[8] + 0.5 * (f[1] - f[3]) + rho * uy / 6 - rho * ux / 2.;
        f[5] = f[7] + 0.5 * (f[3] - f[1]) + rho * uy / 6 + rho * ux / 2.;
        f[2] = f[4] + 2. / 3. * rho * uy;
    } else if constexpr (S == RIGHT) {
        f[6] = f[8] + 0.5 * (f[4] - f[2]) - rho * ux / 6 + rho * uy / 2.;
        f[7] = f[5] + 0.5 * (f[2] - f[4]) - rho * ux / 6 - rho * uy / 2.;
        f[3] = f[1] - 2. / 3. * rho * ux;
    } else {
        f[5] = f[7] + 0.5 * (f[4] - f[2]) + rho * ux / 6. + rho * uy / 2.;
        f[8] = f[6] + 0.5 * (f[2] - f[4]) + rho * ux / 6. - rho * uy / 2.;
        f[1] = f[3] + 2. / 3. * rho * ux;
    }
}

// Non-equilibrium momentum flux from the known populations (xNeqPi,
// boundaries.c:266-302).  n<I>() yields fNeq_I = f_I - feq_I (splitEqNeq,
// boundaries.c:239-249); only the populations a side uses are formed.
template <int S, typename N>
__device__ __forceinline__ void neq_pi(N &&n, double &pxx, double &pyy, double &pxy) {
    FP64_EXACT;
    if constexpr (S == UPPER) {
        double n1 = n.template operator()<1>(), n2 = n.template operator()<2>();
        double n3 = n.template operator()<3>(), n5 = n.template operator()<5>();
        double n6 = n.template operator()<6>();
        pxx = n1 + n3 + 2. * (n5 + n6);
        pyy = 2. * (n6 + n2 + n5);
        pxy = 2. * (n5 - n6);
    } else if constexpr (S == LOWER) {
        double n1 = n.template operator()<1>(), n3 = n.template operator()<3>();
        double n4 = n.template operator()<4>(), n7 = n.template operator()<7>();
        double n8 = n.template operator()<8>();
        pxx = n1 + n3 + 2. * (n7 + n8);
        pyy = 2. * (n4 + n7 + n8);
        pxy = 2. * (n7 - n8);
    } else if constexpr (S == RIGHT) {
        double n1 = n.template operator()<1>(), n2 = n.template operator()<2>();
        double n4 = n.template operator()<4>(), n5 = n.template operator()<5>();
        double n8 = n.template operator()<8>();
        pxx = 2. * (n1 + n5 + n8);
        pyy = n2 + n4 + 2. * (n5 + n8);
        pxy = 2. * (n5 - n8);
    } else {
        double n2 = n.template operator()<2>(), n3 = n.template operator()<3>();
        double n4 = n.template operator()<4>(), n6 = n.template operator()<6>();
        double n7 = n.template operator()<7>();
        pxx = 2. * (n3 + n6 + n7);
        pyy = n2 + n4 + 2. * (n6 + n7);
        pxy = 2. * (n7 - n6);
    }
}

// Regularized reconstruction (splitEqNeq + xNeqPi + regularizedF,
// boundaries.c:239-262): f_i = feq_i + 9/2 w_i ((cx^2-1/3) Pxx +
// (cy^2-1/3) Pyy + 2 cx cy Pxy), the bracket summed left to right.  feq_i is
// evaluated again for the reconstruction (the same operations, the same
// bits) instead of being kept in nine registers.
template <int I>
__device__ __forceinline__ double reg_term(double pxx, double pyy, double pxy) {
    FP64_EXACT;
    constexpr double qx = (double)(cx(I) * cx(I)) - 1. / 3.;
    constexpr double qy = (double)(cy(I) * cy(I)) - 1. / 3.;
    constexpr double qxy = 2. * cx(I) * cy(I);
    return 9. / 2. * wgt(I) * (qx * pxx + qy * pyy + qxy * pxy);
}

template <int S>
__device__ __forceinline__ void regularize(double f[9], double rho, double ux, double uy) {
    FP64_EXACT;
    double usqr = ux * ux + uy * uy;
    auto neq = [&]<int I>() {
        FP64_EXACT;
        return f[I] - feq<I>(rho, ux, uy, usqr);
    };
    double pxx, pyy, pxy;
    neq_pi<S>(neq, pxx, pyy, pxy);
    for_pops([&]<int I>() {
        FP64_EXACT;
        f[I] = feq<I>(rho, ux, uy, usqr) + reg_term<I>(pxx, pyy, pxy);
    });
}

// Wall operator families and kinds of a channel side.
enum BcFamily : int { BC_REGULARIZED = 0, BC_ZOUHE = 1 };
enum BcKind : int { BC_VELOCITY = 0, BC_PRESSURE = 1 };

// One straight-wall node: the operator of (side, kind, family) followed by
// the bulk BGK, as every boundaries.c operator ends (e.g. boundaries.c:314).
// p0/p1 are (ux, uy) for a velocity operator, (rho, uPar) for a pressure one.
template <int S>
__device__ __forceinline__ void wall_node(double f[9], int kind, int family, double p0,
                                          double p1, double omega) {
    FP64_EXACT;
    constexpr bool normal_y = (S == UPPER || S == LOWER);
    if (kind == BC_VELOCITY) {
        double rho = wall_rho<S>(f, normal_y ? p1 : p0);
        if (family == BC_ZOUHE) zh_complete<S>(f, p0, p1, rho);   // boundaries.c:171-201
        else regularize<S>(f, rho, p0, p1);                        // boundaries.c:304-358
    } else {
        double un = wall_un<S>(f, p0);
        double ux = normal_y ? p1 : un, uy = normal_y ? un : p1;
        if (family == BC_ZOUHE) zh_complete<S>(f, ux, uy, p0);    // boundaries.c:203-233
        else regularize<S>(f, p0, ux, uy);                         // boundaries.c:360-414
    }
    bgk(f, omega);
}

}  // namespace lbm
