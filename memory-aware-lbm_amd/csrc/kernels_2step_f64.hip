// kernels_2step_f64.hip -- two time steps per launch, FP64, bit-exact.
//
// The reference's memory-aware sweeps (collide_tight_block, lb.c:936-1069;
// collide_tight_panel_iy, lb.c:1402-1538) do the second collide-and-stream of
// a node as soon as all its inputs from the first one exist, so the lattice
// crosses the memory hierarchy once per two steps.  On the GPU the first
// step's results live in LDS instead of cache:
//
//  k_block  a workgroup owns a 2-D tile; the first step covers the tile plus
//           a one-node ring in LDS, the second step the tile itself
//           (block_tile.hpp, shared with the batched channels).
//  (the panel sweep keeps its first step in wave registers instead:
//   kernels_wave.hip)
#include "launch.hpp"
#include "block_tile.hpp"

namespace lbm {

// ---- 2-D tiles --------------------------------------------------------------

template <int TYP>  // padded tile height (rows incl. the one-node ring), power of two
__global__ __launch_bounds__(BLOCK_THREADS) void k_block(Problem p, BlockGrid g,
                                                         const double *__restrict__ P,
                                                         double *__restrict__ Q) {
    extern __shared__ double lds[];  // [9][tx+2][TYP]
    const int by = blockIdx.x % g.nby;
    const int bx = blockIdx.x / g.nby;
    block_tile<TYP>(p, g.tx, g.nbx, bx, by, P, Q, lds);
}

// ---- launchers ----------------------------------------------------------------

BlockGrid block_grid(const Layout &L, int tx, int typ) {
    BlockGrid g;
    if (typ != 32 && typ != 64 && typ != 128) typ = 64;
    if (tx < 3) tx = 8;   // 46 KB of LDS: three workgroups per CU
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
