// rows.hpp -- the one-step sweeps (fused pull + collide, local collide,
// pure stream gather) with every lane holding R contiguous rows of a column
// per population, R * sizeof(S) bytes loaded and stored at once (up to 16:
// FP64 R = 2, FP32 4, FP16 8 on large lattices).
//
// A wave covers rows y0 .. y0 + 64R - 1 of one column (y0 - 1 a multiple of
// 64R), lane l rows y0 + R l .. y0 + R l + R - 1.  Every population is read
// with one aligned R-row load per lane, whatever its c_y: the pull's one-row
// offset is taken from the lane's own vector and, for the vector's first
// (c_y = +1) or last (c_y = -1) row, from the neighbouring lane by a DPP
// shift; the wave's two outer rows come from one extra load per pair of
// populations sharing a source column (lane 0 reads row y0 - 1 of the c_y =
// +1 population, lane 63 row y0 + 64R of the c_y = -1 one).  Stores are
// aligned non-temporal stores of the same width.  So a wave issues nine
// loads, three short halo loads and nine stores for 64R nodes, each touching
// whole aligned pieces of cache lines, instead of nine loads (six of them
// misaligned by one element) and nine stores per 64 nodes.  A wave covers
// 64R rows of one column, so R shrinks on small lattices to keep enough
// waves in flight (rows_per_lane).
//
// The per-node collision, perimeter bookkeeping and outlet extrapolation are
// the single-node kernels' own (collide_node_f64 / collide_node_lp), so the
// results are the same bits.
#pragma once

#include <type_traits>
#include <utility>

#include "lanes.hpp"
#include "node_step.hpp"

namespace lbm {

enum RowMode : int { ROWS_COLLIDE_LOCAL = 0, ROWS_COLLIDE_PULL = 1, ROWS_STREAM = 2 };

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// R rows of one population as one memory word of R * sizeof(S) bytes
template <int BYTES> struct WordOf;
template <> struct WordOf<16> { using T = u32x4; };
template <> struct WordOf<8> { using T = u32x2; };
template <> struct WordOf<4> { using T = uint32_t; };
template <> struct WordOf<2> { using T = uint16_t; };

template <typename S, int R> struct RowV {
    S v[R];
};

template <typename S, int R> __device__ __forceinline__ RowV<S, R> load_rows(const S *p) {
    using W = typename WordOf<R * (int)sizeof(S)>::T;
    return __builtin_bit_cast(RowV<S, R>, *(const W *)p);
}
template <typename S, int R> __device__ __forceinline__ void store_rows_nt(const RowV<S, R> &r, S *p) {
    using W = typename WordOf<R * (int)sizeof(S)>::T;
    __builtin_nontemporal_store(__builtin_bit_cast(W, r), (W *)p);
}

// fn.template operator()<J>() for J = 0 .. R-1, each with a constant index
// (a loop the compiler unrolls only partly would index the node arrays at
// run time and move them to scratch memory)
template <typename F, int... J>
__device__ __forceinline__ void for_rows_impl(F &&fn, std::integer_sequence<int, J...>) {
    (fn.template operator()<J>(), ...);
}
template <int R, typename F> __device__ __forceinline__ void for_rows(F &&fn) {
    for_rows_impl(fn, std::make_integer_sequence<int, R>{});
}

// Rows per lane for a lattice: the widest (at most 16 bytes per lane and
// population) that still gives 32768 waves, 128 per compute unit (on 1024 x
// 256 FP32, four rows per lane left 1024 waves and ran at 16 GLUPS against
// 29 with one row).
inline int rows_per_lane(const Layout &L, int elem_bytes) {
    int r = 16 / elem_bytes;
    while (r > 1 && (int64_t)L.lx * ((L.ly + 64 * r - 1) / (64 * r)) < 32768) r /= 2;
    return r;
}

template <typename S, int MODE, int R>
__device__ __forceinline__ void rows_body(const Problem &p, const S *__restrict__ src,
                                          S *__restrict__ dst) {
    constexpr bool PULL = MODE != ROWS_COLLIDE_LOCAL;
    constexpr bool F64 = sizeof(S) == 8;
    using T = typename std::conditional<F64, double, float>::type;
    const Layout &L = p.lay;
    const int lane = threadIdx.x & 63;
    const int y0 = 1 + (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 * R);
    if (y0 > L.ly) return;
    const int x = blockIdx.y + 1;
    const int ya = y0 + R * lane;
    // A vector starting at a row <= ly + 1 lies inside its column segment
    // (segments hold rows up to ly + 1 and end on a 64-element boundary);
    // lanes past that read the wave's first vector, and nothing they hold
    // reaches a node <= ly.
    const int yl = ya <= L.ly + 1 ? ya : y0;

    T f[R][9];
    for_pops([&]<int I>() {
        constexpr int sx = PULL ? cx(I) : 0, sy = PULL ? cy(I) : 0;
        const S *col = src + I * L.plane + (int64_t)(x - sx) * L.pitch + L.yoff;
        const RowV<S, R> v = load_rows<S, R>(col + yl);
        if constexpr (sy == 0) {
#pragma unroll
            for (int j = 0; j < R; ++j) f[j][I] = (T)v.v[j];
        } else if constexpr (sy == 1) {
            f[0][I] = (T)lane_up(v.v[R - 1]);
#pragma unroll
            for (int j = 1; j < R; ++j) f[j][I] = (T)v.v[j - 1];
        } else {
#pragma unroll
            for (int j = 0; j < R - 1; ++j) f[j][I] = (T)v.v[j + 1];
            f[R - 1][I] = (T)lane_down(v.v[0]);
        }
    });
    if constexpr (PULL) {
        // the wave's outer rows: lane 0 row y0 - 1 of population A (c_y = +1),
        // lane 63 row y0 + 64R (at most ly + 1) of population B (c_y = -1)
        const int yhi = y0 + 64 * R <= L.ly + 1 ? y0 + 64 * R : L.ly + 1;
        auto halo = [&]<int A, int B>() {
            static_assert(cx(A) == cx(B) && cy(A) == 1 && cy(B) == -1, "halo pair");
            const int64_t col = (int64_t)(x - cx(A)) * L.pitch + L.yoff;
            const S *hp = lane == 63 ? src + B * L.plane + col + yhi
                                     : src + A * L.plane + col + (y0 - 1);
            const T h = (T)*hp;
            if (lane == 0) f[0][A] = h;
            if (lane == 63) f[R - 1][B] = h;
        };
        halo.template operator()<2, 4>();
        halo.template operator()<5, 8>();
        halo.template operator()<6, 7>();
    }

    if constexpr (MODE != ROWS_STREAM) {
        uint32_t step = 0u;
        if constexpr (sizeof(S) == 4) {
            step = *p.step;
            fp32_round_toward_zero();
        }
        for_rows<R>([&]<int J>() {
            if (ya + J <= L.ly) {
                if constexpr (F64) collide_node_f64<PULL>(p, x, ya + J, f[J], src, dst);
                else collide_node_lp<S, PULL, true>(p, x, ya + J, f[J], src, dst, step);
            }
        });
    }

    auto out = [](T v) -> S {
        if constexpr (F64) return v;
        else return narrow<S>(v);
    };
    for_pops([&]<int I>() {
        S *col = dst + I * L.plane + (int64_t)x * L.pitch + L.yoff;
        if (ya + R - 1 <= L.ly) {
            RowV<S, R> o;
#pragma unroll
            for (int j = 0; j < R; ++j) o.v[j] = out(f[j][I]);
            store_rows_nt<S, R>(o, col + ya);
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j)
                if (ya + j <= L.ly) col[ya + j] = out(f[j][I]);
        }
    });
}

// One step of every interior node from a collided state (PULL: the fused
// collide-stream of combine / tight) or from a streamed one (the origin
// sweep's collision, lb.c:172-201), into the other buffer.
template <typename S, bool PULL, int R>
__global__ __launch_bounds__(256) void k_collide(Problem p, const S *__restrict__ src,
                                                 S *__restrict__ dst) {
    rows_body<S, PULL ? ROWS_COLLIDE_PULL : ROWS_COLLIDE_LOCAL, R>(p, src, dst);
}

// The pure stream of a collided state (the origin sweep's propagation,
// lb.c:246-287, as a pull gather); only p.lay is read.
template <typename S, int R>
__global__ __launch_bounds__(256) void k_stream(Problem p, const S *__restrict__ src,
                                                S *__restrict__ dst) {
    rows_body<S, ROWS_STREAM, R>(p, src, dst);
}

// Grid of k_collide / k_stream at R rows per lane: up to four waves per
// workgroup along the column, one column per grid row.
inline dim3 rows_grid(const Layout &L, int R, int &threads) {
    const int waves = (L.ly + 64 * R - 1) / (64 * R);
    const int wpb = waves < 4 ? waves : 4;
    threads = 64 * wpb;
    return dim3((waves + wpb - 1) / wpb, L.lx);
}

// k_collide (mode 0 / 1: local / pull) or k_stream (mode 2) at the rows per
// lane the lattice takes
template <typename S>
inline hipError_t launch_rows(int mode, const Problem &p, const S *src, S *dst, hipStream_t st) {
    const int R = rows_per_lane(p.lay, (int)sizeof(S));
    int t;
    const dim3 g = rows_grid(p.lay, R, t);
    auto go = [&](auto rt) {
        constexpr int RR = decltype(rt)::value;
        if constexpr (RR * sizeof(S) <= 16) {
            if (mode == 1) k_collide<S, true, RR><<<g, t, 0, st>>>(p, src, dst);
            else if (mode == 0) k_collide<S, false, RR><<<g, t, 0, st>>>(p, src, dst);
            else k_stream<S, RR><<<g, t, 0, st>>>(p, src, dst);
        }
    };
    switch (R) {
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    default: go(std::integral_constant<int, 1>{}); break;
    }
    return hipGetLastError();
}

}  // namespace lbm
