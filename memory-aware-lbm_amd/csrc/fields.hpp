// fields.hpp -- moments and reductions of the post-stream state, for every
// storage format (FP64 populations, or FP32/FP16 deviations from the weights
// which are widened to FP64 on load).
#pragma once

#include "collide_f64.hpp"

namespace lbm {

// population i as an FP64 value
template <typename S, int I> __device__ __forceinline__ double pop_f64(const S *p) {
    if constexpr (sizeof(S) == 8) return *p;
    else return wgt(I) + (double)(float)*p;
}

template <typename S, bool PULL>
__device__ __forceinline__ void load_post_stream(const Layout &L, const S *__restrict__ src,
                                                 int x, int y, double f[9]) {
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() {
        f[I] = PULL ? pop_f64<S, I>(&src[I * L.plane + n - cx(I) * L.pitch - cy(I)])
                    : pop_f64<S, I>(&src[I * L.plane + n]);
    });
}

// computeMacros (lb.c:134-141) on every interior node of columns x0 ..
// x0 + gridDim.y - 1, x-major output from column x0
template <typename S, bool PULL>
__global__ void k_macros(Layout L, int x0, const S *__restrict__ src, double *rho, double *ux,
                         double *uy) {
    int y = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int x = blockIdx.y + x0;
    if (y > L.ly) return;
    double f[9];
    load_post_stream<S, PULL>(L, src, x, y, f);
    double r, u, v;
    macros(f, r, u, v);
    int64_t o = (int64_t)(x - x0) * L.ly + (y - 1);
    if (rho) rho[o] = r;
    if (ux) ux[o] = u;
    if (uy) uy[o] = v;
}

__device__ __forceinline__ bool solid_node(const Problem &p, int x, int y) {
    return p.r2 >= 0 && side_of(p, x, y) < 0 && in_disc(p, x, y);
}

// drag/lift by momentum exchange, kinetic energy and mass: block partial
// sums in a fixed order (wave64 tree, then the block's waves)
template <typename S, bool PULL>
__global__ __launch_bounds__(256) void k_diag(Problem p, const S *__restrict__ src,
                                              double *__restrict__ partial) {
    FP64_EXACT;
    const Layout &L = p.lay;
    int y = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int x = blockIdx.y + 1;
    double acc[4] = {0., 0., 0., 0.};
    if (y <= L.ly) {
        double f[9];
        load_post_stream<S, PULL>(L, src, x, y, f);
        bool solid = solid_node(p, x, y);
        if (!solid) {
            FP64_EXACT;
            double r, u, v;
            macros(f, r, u, v);
            acc[2] = 0.5 * r * (u * u + v * v);
            acc[3] = r;
        }
        for_pops([&]<int I>() {
            if constexpr (I != 0) {
                int sx = x - cx(I), sy = y - cy(I);
                if (!outside(L, sx, sy)) {
                    bool from_solid = solid_node(p, sx, sy);
                    if (solid && !from_solid) { acc[0] += cx(I) * f[I]; acc[1] += cy(I) * f[I]; }
                    if (!solid && from_solid) { acc[0] -= cx(I) * f[I]; acc[1] -= cy(I) * f[I]; }
                }
            }
        });
    }
    __shared__ double red[4][4];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 4; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        int nw = (blockDim.x + 63) >> 6;
        double v = 0.;
        for (int w = 0; w < nw; ++w) v += red[w][threadIdx.x];
        int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
        partial[b * 4 + threadIdx.x] = v;
    }
}

}  // namespace lbm
