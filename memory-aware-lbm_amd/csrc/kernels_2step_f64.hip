// kernels_2step_f64.hip -- two time steps per launch through LDS tiles
// (bit-exact FP64, and the FP32 / FP16 storage forms).
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

template <typename S, int TYP>  // TYP: padded tile height (rows + ring), power of two
__global__ __launch_bounds__(BLOCK_THREADS) void k_block(Problem p, BlockGrid g,
                                                         const S *__restrict__ P,
                                                         S *__restrict__ Q) {
    extern __shared__ __align__(16) unsigned char lds_raw[];  // [9][tx+2][TYP] compute-type values
    int by = blockIdx.x % g.nby;
    int bx = blockIdx.x / g.nby;
    if constexpr (sizeof(S) == 4) {
        pin_index(by);
        pin_index(bx);
    }
    const uint32_t m = sizeof(S) == 4 ? *p.step : 0u;   // the launch's first collision
    block_tile<S, TYP>(p, g.tx, g.nbx, bx, by, m, P, Q, (typename Calc<S>::T *)lds_raw);
}

// ---- launchers ----------------------------------------------------------------

BlockGrid block_grid(const Layout &L, int tx, int typ, int tbytes) {
    BlockGrid g;
    if (typ != 32 && typ != 64 && typ != 128) typ = 64;
    // default (fastest on 16384 x 4096): FP64 8 columns, 46 KB of LDS and
    // three workgroups per CU; FP32 / FP16 values 14 columns, 37 KB, four
    if (tx < 3) tx = tbytes == 8 ? 8 : 14;
    // LDS: 9 populations x (tx+2) x typ values must fit 160 KB
    while (tx > 3 && 9 * (size_t)(tx + 2) * typ * tbytes > 160 * 1024) --tx;
    g.tx = tx;
    g.typ = typ;
    g.tbytes = tbytes;
    g.nby = (L.ly + (typ - 2) - 1) / (typ - 2);
    g.nbx = (L.lx + tx - 1) / tx;
    return g;
}

size_t block_lds_bytes(const BlockGrid &g) { return 9 * (size_t)(g.tx + 2) * g.typ * g.tbytes; }

template <typename S>
static hipError_t block_go(const Problem &p, const BlockGrid &g, const void *P, void *Q,
                           hipStream_t st) {
    const size_t lds = block_lds_bytes(g);
    if (lds > 160 * 1024 || g.tx < 3 || g.tbytes != (int)sizeof(typename Calc<S>::T))
        return hipErrorInvalidValue;
    dim3 grid(g.nbx * g.nby);
    const S *src = (const S *)P;
    S *dst = (S *)Q;
    switch (g.typ) {
    case 32: k_block<S, 32><<<grid, BLOCK_THREADS, lds, st>>>(p, g, src, dst); break;
    case 64: k_block<S, 64><<<grid, BLOCK_THREADS, lds, st>>>(p, g, src, dst); break;
    case 128: k_block<S, 128><<<grid, BLOCK_THREADS, lds, st>>>(p, g, src, dst); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_block(int storage, const Problem &p, const BlockGrid &g, const void *P, void *Q,
                        hipStream_t st) {
    switch (storage) {
    case 0: return block_go<double>(p, g, P, Q, st);
    case 1: return block_go<float>(p, g, P, Q, st);
    default: return block_go<_Float16>(p, g, P, Q, st);
    }
}

template <typename S> static hipError_t allow_lds() {
    // up to 160 KB of dynamic LDS per workgroup
    hipError_t e = hipSuccess;
    for (const void *k : {(const void *)k_block<S, 32>, (const void *)k_block<S, 64>,
                          (const void *)k_block<S, 128>})
        if (e == hipSuccess)
            e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    return e;
}

hipError_t block_kernel_setup() {
    hipError_t e = allow_lds<double>();
    if (e == hipSuccess) e = allow_lds<float>();
    if (e == hipSuccess) e = allow_lds<_Float16>();
    return e;
}

}  // namespace lbm
