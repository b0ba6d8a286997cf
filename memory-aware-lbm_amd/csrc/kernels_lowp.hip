// kernels_lowp.hip -- reduced-precision storage: FP32 or FP16 populations.
//
// What is stored is the deviation g_i = f_i - w_i from the rest-state
// weights: it is O(u) instead of O(1), so the storage mantissa is spent on the
// flow rather than on the constant part.  Bulk collisions run in FP32 on the
// deviations:
//   rho = 1 + sum g,  j = sum c g,  u = j / rho,
//   geq_i = w_i (drho + rho (3 c.u + 9/2 (c.u)^2 - 3/2 u^2)),
//   g_i  <- g_i + omega (geq_i - g_i).
// Perimeter nodes widen to FP64 (f = w + g) and run the FP64 wall operators of
// the bit-exact path, including the outlet densities of the zero-gradient
// extrapolation; bounce-back is a permutation.  Moments and reductions are
// accumulated in FP64.  The state conventions (streamed / collided, the ghost
// ring carrying the unknown populations) are those of kernels_f64.hip.
//
// No counterpart in the reference (its arithmetic is FP64 throughout); the
// error against the FP64 path is measured, not assumed (tests/test_gpu_precision.py).
#include "launch.hpp"
#include "collide_f64.hpp"
#include "fields.hpp"
#include "lowp_ops.hpp"
#include "node_step.hpp"
#include "rows.hpp"

namespace lbm {

template <typename S>
__global__ void k_init_lp(Layout L, const double *__restrict__ init_ux, S *__restrict__ dst) {
    int y = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int x = blockIdx.y + 1;
    if (y > L.ly) return;
    double u = init_ux[y];
    double usqr = u * u;
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() { dst[I * L.plane + n] = (S)(feq<I>(1., u, 0., usqr) - wgt(I)); });
}

// f = 0 everywhere (the reference's calloc'd buffers) in deviation storage:
// Flat kernels over a whole buffer of `span` elements; the population of
// element k is (k / plane) % 9 (planes: plane = one population's elements;
// interleaved columns: plane = one column's segment of one population).
__device__ __forceinline__ double w_of(int i) { return i == 0 ? 4. / 9. : (i <= 4 ? 1. / 9. : 1. / 36.); }

template <typename S>
__global__ void k_fill_zero_f(int64_t span, int64_t plane, S *__restrict__ dst) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= span) return;
    dst[k] = (S)(-w_of((int)((k / plane) % 9)));
}

template <typename S>
__global__ void k_widen(int64_t span, int64_t plane, const S *__restrict__ src, double *__restrict__ dst) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= span) return;
    dst[k] = w_of((int)((k / plane) % 9)) + (double)(float)src[k];
}

template <typename S>
__global__ void k_narrow(int64_t span, int64_t plane, const double *__restrict__ src, S *__restrict__ dst) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= span) return;
    dst[k] = (S)(src[k] - w_of((int)((k / plane) % 9)));
}

// ---- launchers ----------------------------------------------------------

static dim3 grid_lp(const Layout &L, int &threads) {
    threads = L.ly >= 256 ? 256 : ((L.ly + 63) / 64) * 64;
    return dim3((L.ly + threads - 1) / threads, L.lx);
}

template <typename S>
static hipError_t collide_t(const Problem &p, bool pull, const void *src, void *dst, hipStream_t st) {
    return launch_rows<S>(pull ? 1 : 0, p, (const S *)src, (S *)dst, st);
}

template <typename S>
static hipError_t stream_t(const Layout &L, const void *src, void *dst, hipStream_t st) {
    Problem q{};
    q.lay = L;
    return launch_rows<S>(2, q, (const S *)src, (S *)dst, st);
}

hipError_t launch_collide_lp(int storage, const Problem &p, bool pull, const void *src, void *dst,
                             hipStream_t st) {
    return storage == 1 ? collide_t<float>(p, pull, src, dst, st)
                        : collide_t<_Float16>(p, pull, src, dst, st);
}

hipError_t launch_stream_lp(int storage, const Layout &L, const void *src, void *dst,
                            hipStream_t st) {
    return storage == 1 ? stream_t<float>(L, src, dst, st) : stream_t<_Float16>(L, src, dst, st);
}

hipError_t launch_init_lp(int storage, const Layout &L, const double *init_ux, void *dst,
                          hipStream_t st) {
    int t;
    dim3 g = grid_lp(L, t);
    if (storage == 1) k_init_lp<float><<<g, t, 0, st>>>(L, init_ux, (float *)dst);
    else k_init_lp<_Float16><<<g, t, 0, st>>>(L, init_ux, (_Float16 *)dst);
    return hipGetLastError();
}

hipError_t launch_macros_lp(int storage, const Layout &L, int x0, int nx, bool pull,
                            const void *src, double *rho, double *ux, double *uy, hipStream_t st) {
    int t;
    dim3 g = grid_lp(L, t);
    g.y = nx;
    if (storage == 1) {
        const float *s = (const float *)src;
        if (pull) k_macros<float, true><<<g, t, 0, st>>>(L, x0, s, rho, ux, uy);
        else k_macros<float, false><<<g, t, 0, st>>>(L, x0, s, rho, ux, uy);
    } else {
        const _Float16 *s = (const _Float16 *)src;
        if (pull) k_macros<_Float16, true><<<g, t, 0, st>>>(L, x0, s, rho, ux, uy);
        else k_macros<_Float16, false><<<g, t, 0, st>>>(L, x0, s, rho, ux, uy);
    }
    return hipGetLastError();
}

hipError_t launch_diag_lp(int storage, const Problem &p, bool pull, const void *src,
                          double *partial, hipStream_t st) {
    int t;
    dim3 g = grid_lp(p.lay, t);
    if (storage == 1) {
        if (pull) k_diag<float, true><<<g, t, 0, st>>>(p, (const float *)src, partial);
        else k_diag<float, false><<<g, t, 0, st>>>(p, (const float *)src, partial);
    } else {
        if (pull) k_diag<_Float16, true><<<g, t, 0, st>>>(p, (const _Float16 *)src, partial);
        else k_diag<_Float16, false><<<g, t, 0, st>>>(p, (const _Float16 *)src, partial);
    }
    return hipGetLastError();
}

hipError_t launch_fill_zero_f(int storage, const Layout &L, void *dst, hipStream_t st) {
    const int64_t span = L.span();
    const unsigned g = (unsigned)((span + 255) / 256);
    if (storage == 1) k_fill_zero_f<float><<<g, 256, 0, st>>>(span, L.plane, (float *)dst);
    else k_fill_zero_f<_Float16><<<g, 256, 0, st>>>(span, L.plane, (_Float16 *)dst);
    return hipGetLastError();
}

hipError_t launch_widen(int storage, const Layout &L, const void *src, double *dst, hipStream_t st) {
    const int64_t span = L.span();
    const unsigned g = (unsigned)((span + 255) / 256);
    if (storage == 1) k_widen<float><<<g, 256, 0, st>>>(span, L.plane, (const float *)src, dst);
    else k_widen<_Float16><<<g, 256, 0, st>>>(span, L.plane, (const _Float16 *)src, dst);
    return hipGetLastError();
}

hipError_t launch_narrow(int storage, const Layout &L, const double *src, void *dst, hipStream_t st) {
    const int64_t span = L.span();
    const unsigned g = (unsigned)((span + 255) / 256);
    if (storage == 1) k_narrow<float><<<g, 256, 0, st>>>(span, L.plane, src, (float *)dst);
    else k_narrow<_Float16><<<g, 256, 0, st>>>(span, L.plane, src, (_Float16 *)dst);
    return hipGetLastError();
}

}  // namespace lbm
