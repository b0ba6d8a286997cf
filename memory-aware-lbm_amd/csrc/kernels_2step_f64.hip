// kernels_2step_f64.hip -- two time steps per launch, FP64, bit-exact.
//
// The reference's memory-aware sweeps (collide_tight_block, lb.c:936-1069;
// collide_tight_panel_iy, lb.c:1402-1538) do the second collide-and-stream of
// a node as soon as all its inputs from the first one exist, so the lattice
// crosses the memory hierarchy once per two steps.  On the GPU the first
// step's results live in LDS instead of cache:
//
//  k_block  a workgroup owns a 2-D tile; the first step covers the tile plus
//           a one-node ring in LDS, the second step the tile itself.
//  (the panel sweep keeps its first step in wave registers instead:
//   kernels_wave.hip)
//
// k_block reads only the collided buffer P (post-collision of step m-1, ghost
// ring = unknown populations for step m) and write only Q (post-collision of
// step m+1 and the ghost ring for step m+2), so workgroups never race.  The
// first step's unknown populations come from P's ghost ring; the second
// step's are the node's own post-collision populations of step m-1, i.e.
// P(x) itself.  The outlet density of the second step is extrapolated from
// the densities of columns lx-2 and lx-1 taken just before their second
// collision, as the reference captures them (lb.c:968-977, 1025-1029);
// every launch therefore gives the tile that holds column lx at least three
// columns (tiles are aligned to the right end).
#include "launch.hpp"
#include "collide_f64.hpp"

namespace lbm {

// LDS-only workgroup barrier: orders the tile's LDS traffic without waiting
// for outstanding global loads (a plain __syncthreads() would drain them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ void pull_node(const Layout &L, const double *__restrict__ P, int x,
                                          int y, double f[9]) {
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() { f[I] = P[I * L.plane + n - cx(I) * L.pitch - cy(I)]; });
}

__device__ __forceinline__ double pulled_density(const Layout &L, const double *__restrict__ P,
                                                 int x, int y) {
    double g[9];
    pull_node(L, P, x, y, g);
    return rho_of(g);
}

// First step of node (x, y): post-stream from P (ghost ring included), then
// the node's collision with the outlet density of this step.
__device__ __forceinline__ void first_step(const Problem &p, const double *__restrict__ P, int x,
                                           int y, double f[9]) {
    int s = side_of(p, x, y);
    double rp0 = 0.;
    if (s == RIGHT && p.zgb)
        rp0 = zgb_rho(pulled_density(p.lay, P, p.lay.lx - 1, y),
                      pulled_density(p.lay, P, p.lay.lx - 2, y));
    collide_class(p, s, x, y, f, rp0);
}

// ---- 2-D tiles --------------------------------------------------------------

constexpr int BLOCK_THREADS = 256;

template <int TYP>  // padded tile height (rows incl. the one-node ring), power of two
__global__ __launch_bounds__(BLOCK_THREADS) void k_block(Problem p, BlockGrid g,
                                                         const double *__restrict__ P,
                                                         double *__restrict__ Q) {
    extern __shared__ double lds[];  // [9][TXP][TYP]
    const Layout &L = p.lay;
    const int TXP = g.tx + 2;
    const int area = TXP * TYP;
    const int by = blockIdx.x % g.nby;
    const int bx = blockIdx.x / g.nby;
    const int Y0 = 1 + by * (TYP - 2);
    const int X1 = L.lx + 1 - (g.nbx - 1 - bx) * g.tx;
    const int X0 = X1 - g.tx > 1 ? X1 - g.tx : 1;
    const int Y1 = Y0 + TYP - 2 < L.ly + 1 ? Y0 + TYP - 2 : L.ly + 1;

    // step m on the tile plus its ring
    for (int k = threadIdx.x; k < area; k += BLOCK_THREADS) {
        int a = k / TYP, b = k % TYP;
        int x = X0 - 1 + a, y = Y0 - 1 + b;
        if (x > X1 || outside(L, x, y)) continue;
        double f[9];
        pull_node(L, P, x, y, f);
        first_step(p, P, x, y, f);
        for_pops([&]<int I>() { lds[I * area + k] = f[I]; });
    }
    lds_barrier();
    // step m+1 on the tile
    for (int k = threadIdx.x; k < area; k += BLOCK_THREADS) {
        int a = k / TYP, b = k % TYP;
        int x = X0 - 1 + a, y = Y0 - 1 + b;
        if (x < X0 || x >= X1 || y < Y0 || y >= Y1) continue;
        int64_t n = L.node(x, y);
        double h[9];
        for_pops([&]<int I>() {
            int sx = x - cx(I), sy = y - cy(I);
            h[I] = outside(L, sx, sy) ? P[I * L.plane + n] : lds[I * area + k - cx(I) * TYP - cy(I)];
        });
        int s = side_of(p, x, y);
        double rp0 = 0.;
        if (s >= 0) {
            for_pops([&]<int I>() {
                if constexpr (I != 0) {
                    int gx = x - cx(I), gy = y - cy(I);
                    if (outside(L, gx, gy)) Q[I * L.plane + L.node(gx, gy)] = lds[I * area + k];
                }
            });
            if (s == RIGHT) {
                // post-stream densities of step m+1 at columns lx-1, lx-2, from LDS
                double g1[9], g2[9];
                for_pops([&]<int I>() {
                    g1[I] = lds[I * area + k - TYP - cx(I) * TYP - cy(I)];
                    g2[I] = lds[I * area + k - 2 * TYP - cx(I) * TYP - cy(I)];
                });
                rp0 = zgb_rho(rho_of(g1), rho_of(g2));
                p.outlet_rho[y] = p.zgb ? rp0 : p.side[RIGHT].p0[y];
            }
        }
        collide_class(p, s, x, y, h, rp0);
        for_pops([&]<int I>() { __builtin_nontemporal_store(h[I], &Q[I * L.plane + n]); });
    }
}

// ---- launchers ----------------------------------------------------------------

BlockGrid block_grid(const Layout &L, int tx, int typ) {
    BlockGrid g;
    if (typ != 32 && typ != 64 && typ != 128) typ = 64;
    if (tx < 3) tx = 14;
    // LDS: 9 populations x (tx+2) x typ doubles must fit 160 KB
    while (tx > 3 && 9 * (size_t)(tx + 2) * typ * 8 > 160 * 1024) --tx;
    g.tx = tx;
    g.typ = typ;
    g.nby = (L.ly + (typ - 2) - 1) / (typ - 2);
    g.nbx = (L.lx + tx - 1) / tx;
    return g;
}

size_t block_lds_bytes(const BlockGrid &g) { return 9 * (size_t)(g.tx + 2) * g.typ * 8; }

hipError_t launch_block_f64(const Problem &p, const BlockGrid &g, const double *P, double *Q,
                            hipStream_t st) {
    size_t lds = block_lds_bytes(g);
    if (lds > 160 * 1024 || g.tx < 3) return hipErrorInvalidValue;
    dim3 grid(g.nbx * g.nby);
    switch (g.typ) {
    case 32: k_block<32><<<grid, BLOCK_THREADS, lds, st>>>(p, g, P, Q); break;
    case 64: k_block<64><<<grid, BLOCK_THREADS, lds, st>>>(p, g, P, Q); break;
    case 128: k_block<128><<<grid, BLOCK_THREADS, lds, st>>>(p, g, P, Q); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t block_kernel_setup() {
    // allow up to 160 KB of dynamic LDS
    hipError_t e = hipFuncSetAttribute((const void *)k_block<32>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_block<64>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_block<128>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    return e;
}

}  // namespace lbm
