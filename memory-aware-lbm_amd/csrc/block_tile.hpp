// block_tile.hpp -- the two-step LDS tile, shared by the single-lattice
// kernel (k_block, kernels_2step_f64.hip) and the batched one (k_batch_block,
// batch.hip), so both run the same node operations in the same order, for
// every storage format.
//
// A 256-thread workgroup owns a tile of tx columns x (TYP-2) rows.  Step m
// covers the tile plus a one-node ring and goes to LDS (compute-type values,
// rounded as a store would round them); step m+1 covers the tile from LDS.
// The tile reads only the collided buffer P (post-collision of step m-1,
// ghost ring = unknown populations for step m) and writes only Q
// (post-collision of step m+1 and the ghost ring for step m+2), so tiles
// never race.  The first step's unknown populations come from P's ghost
// ring; the second step's are the node's own post-collision populations of
// step m-1, i.e. P(x) itself.  The outlet density of the second step is
// extrapolated from the densities of columns lx-2 and lx-1 taken just before
// their second collision, as the reference captures them (lb.c:968-977,
// 1025-1029), so the tile that holds column lx needs at least three columns
// (tiles are aligned to the right end) and the channel at least four.
#pragma once

#include "step_any.hpp"

namespace lbm {

constexpr int BLOCK_THREADS = 256;

// LDS-only workgroup barrier: orders the tile's LDS traffic without waiting
// for outstanding global loads (a plain __syncthreads() would drain them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Tile (bx, by) of a lattice cut into nbx right-aligned column strips of tx
// columns and rows of TYP-2; lds holds 9 x (tx+2) x TYP compute-type values;
// m is the index of the tile's first collision (FP32 rounding key).
template <typename S, int TYP>  // TYP: padded tile height (rows + ring), power of two
__device__ __forceinline__ void block_tile(const Problem &p, int tx, int nbx, int bx, int by,
                                           uint32_t m, const S *__restrict__ P,
                                           S *__restrict__ Q, typename Calc<S>::T *lds) {
    using T = typename Calc<S>::T;
    const Layout &L = p.lay;
    const int TXP = tx + 2;
    const int area = TXP * TYP;
    const int Y0 = 1 + by * (TYP - 2);
    const int X1 = L.lx + 1 - (nbx - 1 - bx) * tx;
    const int X0 = X1 - tx > 1 ? X1 - tx : 1;
    const int Y1 = Y0 + TYP - 2 < L.ly + 1 ? Y0 + TYP - 2 : L.ly + 1;
    if constexpr (sizeof(S) == 4) fp32_round_toward_zero();   // no integer division below

    // step m on the tile plus its ring
    for (int k = threadIdx.x; k < area; k += BLOCK_THREADS) {
        int a = k / TYP, b = k % TYP;
        int x = X0 - 1 + a, y = Y0 - 1 + b;
        if (x > X1 || outside(L, x, y)) continue;
        const int64_t n = L.node(x, y);
        T f[9];
        for_pops([&]<int I>() { f[I] = as_calc<S>(P[I * L.plane + n - cx(I) * L.pitch - cy(I)]); });
        const int s = side_of(p, x, y);
        double rp0 = 0.;
        if (s == RIGHT && p.zgb)
            rp0 = zgb_rho(pulled_rho<S>(L, P, L.lx - 1, y), pulled_rho<S>(L, P, L.lx - 2, y));
        collide_any<S>(p, s, x, y, f, rp0, m);
        for_pops([&]<int I>() { lds[I * area + k] = stored<S>(f[I]); });
    }
    lds_barrier();
    // step m+1 on the tile
    for (int k = threadIdx.x; k < area; k += BLOCK_THREADS) {
        int a = k / TYP, b = k % TYP;
        int x = X0 - 1 + a, y = Y0 - 1 + b;
        if (x < X0 || x >= X1 || y < Y0 || y >= Y1) continue;
        const int64_t n = L.node(x, y);
        T h[9];
        for_pops([&]<int I>() {
            int sx = x - cx(I), sy = y - cy(I);
            h[I] = outside(L, sx, sy) ? as_calc<S>(P[I * L.plane + n])
                                      : lds[I * area + k - cx(I) * TYP - cy(I)];
        });
        const int s = side_of(p, x, y);
        double rp0 = 0.;
        if (s >= 0) {
            for_pops([&]<int I>() {
                if constexpr (I != 0) {
                    int gx = x - cx(I), gy = y - cy(I);
                    if (outside(L, gx, gy))
                        Q[I * L.plane + L.node(gx, gy)] = to_storage<S>(lds[I * area + k]);
                }
            });
            if (s == RIGHT) {
                // post-stream densities of step m+1 at columns lx-1, lx-2, from LDS
                T g1[9], g2[9];
                for_pops([&]<int I>() {
                    g1[I] = lds[I * area + k - TYP - cx(I) * TYP - cy(I)];
                    g2[I] = lds[I * area + k - 2 * TYP - cx(I) * TYP - cy(I)];
                });
                rp0 = zgb_rho(rho_of_any<S>(g1), rho_of_any<S>(g2));
                p.outlet_rho[y] = p.zgb ? rp0 : p.side[RIGHT].p0[y];
            }
        }
        collide_any<S>(p, s, x, y, h, rp0, m + 1);
        for_pops([&]<int I>() { __builtin_nontemporal_store(to_storage<S>(h[I]), &Q[I * L.plane + n]); });
    }
}

}  // namespace lbm
