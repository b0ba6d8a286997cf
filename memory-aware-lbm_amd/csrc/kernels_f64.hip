// kernels_f64.hip -- FP64 collide/stream kernels for gfx950 (compiled with
// -ffp-contract=off: the results are bit-identical to the reference C code).
//
// State conventions (both hold the reference's step semantics exactly):
//  * streamed: buffer L holds the post-stream, pre-collision populations of
//    step m on every interior node -- including, on perimeter nodes, the
//    "unknown" populations whose source lies outside the domain; those hold
//    whatever the reference's push left there: the node's own post-collision
//    value of step m-2 (lb.c:258-271 never writes them).  Buffer T holds the
//    post-collision populations of step m-1.  This is exactly the reference's
//    (Simulation.lattice, Simulation.tmpLattice) pair after a step.
//  * collided: buffer P holds the post-collision populations of step m-1 on
//    interior nodes, and its ghost ring holds the unknown populations of the
//    perimeter nodes for step m: ghost (x - c_i) slot i = post-collision of
//    node x at step m-2.  A pull gather  f_i(x) = P(x - c_i)_i  then yields
//    the reference's post-stream populations of step m on every node,
//    perimeter included, with no branch.
//
// Kernels:
//  k_collide<double, false, R>  streamed -> collided: reads L locally, collides,
//                    writes the other buffer (one step's collision).
//  k_collide<double, true, R>   collided -> collided: pull gather + collide (one
//                    fused step), the combine/tight hot kernel.
//  k_stream<double, R>  collided -> streamed: pull gather only.
//  (these three: rows.hpp, R rows of a column per lane and population)
//  k_init            equilibrium initial condition.
//  k_macros / k_diag moments and reductions of the post-stream state.
//  k_periodic        the periodic ghost copy on a streamed state (any storage).
#include "launch.hpp"
#include "collide_f64.hpp"
#include "fields.hpp"
#include "node_step.hpp"
#include "rows.hpp"

namespace lbm {

// iniEquilibrium(node, 1, init_ux[y], 0) on every interior node
// (unsteady.c:164-170, lb.c:71-78).
__global__ void k_init(Layout L, const double *__restrict__ init_ux, double *__restrict__ dst) {
    int y = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int x = blockIdx.y + 1;
    if (y > L.ly) return;
    double u = init_ux[y];
    double usqr = u * u + 0. * 0.;
    int64_t n = L.node(x, y);
    for_pops([&]<int I>() { dst[I * L.plane + n] = feq<I>(1., u, 0., usqr); });
}

__global__ void k_diag_final(const double *__restrict__ partial, int64_t nblocks,
                             double *__restrict__ out) {
    // one wave per quantity; lane-strided partial sums then a tree
    int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double v = 0.;
    for (int64_t b = lane; b < nblocks; b += 64) v += partial[b * 4 + k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) out[k] = v;
}

// ---- node-operator known answers -----------------------------------------

__global__ void k_node_op(int which, int n, double *f, const double *p0, const double *p1,
                          double omega) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double g[9];
    for (int i = 0; i < 9; ++i) g[i] = f[k * 9 + i];
    if (which == 0) bounce_back(g);
    else if (which == 17) bgk(g, omega);
    else {
        // 1..8 Zou/He, 9..16 regularized; within each: 4 velocity then 4
        // pressure operators, sides in the order upper, lower, left, right.
        int fam = which >= 9 ? BC_REGULARIZED : BC_ZOUHE;
        int r = (which - 1) % 8;
        int kind = r >= 4 ? BC_PRESSURE : BC_VELOCITY;
        switch (r & 3) {
        case UPPER: wall_node<UPPER>(g, kind, fam, p0[k], p1[k], omega); break;
        case LOWER: wall_node<LOWER>(g, kind, fam, p0[k], p1[k], omega); break;
        case LEFT: wall_node<LEFT>(g, kind, fam, p0[k], p1[k], omega); break;
        default: wall_node<RIGHT>(g, kind, fam, p0[k], p1[k], omega); break;
        }
    }
    for (int i = 0; i < 9; ++i) f[k * 9 + i] = g[i];
}

// ---- plain copy: the bandwidth yardstick for the roofline ------------------
// One 16-byte element per thread over an exact grid, non-temporal store: the
// fastest plain copy measured on MI355X (6.29 TB/s on 2 and 8 GiB, against
// 5.96 for a grid-stride loop over 262144 workgroups and 5.3-5.7 for two to
// eight elements per thread).

typedef double dbl2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_copy(const dbl2 *__restrict__ src,
                                              dbl2 *__restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(src[i], &dst[i]);
}

// ---- launchers -----------------------------------------------------------

static dim3 node_grid(const Layout &L, int &threads) {
    threads = L.ly >= 256 ? 256 : ((L.ly + 63) / 64) * 64;
    return dim3((L.ly + threads - 1) / threads, L.lx);
}

hipError_t launch_collide_f64(const Problem &p, bool pull, const double *src, double *dst,
                              hipStream_t st) {
    return launch_rows<double>(pull ? 1 : 0, p, src, dst, st);
}

hipError_t launch_stream_f64(const Layout &L, const double *src, double *dst, hipStream_t st) {
    Problem q{};
    q.lay = L;
    return launch_rows<double>(2, q, src, dst, st);
}

int one_step_rows(const Layout &L, int storage) {
    return rows_per_lane(L, storage == 0 ? 8 : storage == 1 ? 4 : 2);
}

hipError_t launch_init_f64(const Layout &L, const double *init_ux, double *dst, hipStream_t st) {
    int t;
    dim3 g = node_grid(L, t);
    k_init<<<g, t, 0, st>>>(L, init_ux, dst);
    return hipGetLastError();
}

hipError_t launch_macros_f64(const Layout &L, int x0, int nx, bool pull, const double *src,
                             double *rho, double *ux, double *uy, hipStream_t st) {
    int t;
    dim3 g = node_grid(L, t);
    g.y = nx;
    if (pull) k_macros<double, true><<<g, t, 0, st>>>(L, x0, src, rho, ux, uy);
    else k_macros<double, false><<<g, t, 0, st>>>(L, x0, src, rho, ux, uy);
    return hipGetLastError();
}

int64_t diag_partials_needed(const Layout &L) {
    int t;
    dim3 g = node_grid(L, t);
    return (int64_t)g.x * g.y;
}

hipError_t launch_diag_f64(const Problem &p, bool pull, const double *src, double *partial,
                           double *out, hipStream_t st) {
    int t;
    dim3 g = node_grid(p.lay, t);
    if (pull) k_diag<double, true><<<g, t, 0, st>>>(p, src, partial);
    else k_diag<double, false><<<g, t, 0, st>>>(p, src, partial);
    k_diag_final<<<1, 256, 0, st>>>(partial, (int64_t)g.x * g.y, out);
    return hipGetLastError();
}

// collision counter of a context or batch: advanced by its own one-thread
// kernel on the stream (so captured graphs advance it on every replay)
__global__ void k_bump(uint32_t *ctr, uint32_t by) {
    if (threadIdx.x == 0) *ctr += by;
}

hipError_t launch_bump(uint32_t *ctr, uint32_t by, hipStream_t st) {
    k_bump<<<1, 64, 0, st>>>(ctr, by);
    return hipGetLastError();
}

// Periodic ghost copy (makePeriodic, lb.c:1783-1817) on a streamed state:
// each perimeter node receives, in the slots of populations arriving from
// outside, what left the domain through the opposite side.  The reference
// copies those values from its ghost ring, where propagate (lb.c:258-267)
// pushed them; here they are read at their origin in T, the post-collision
// state the push came from: ghost slot i of (gx, gy) holds T_i(gx - c_i,
// gy - c_i) when that node is interior, and 0 (the calloc'd buffer, in
// deviation storage -w_i) when no interior node feeds it.  The row pass, the
// column pass and the four corner diagonals write disjoint slots once the
// four corner slots, which the reference writes three times, are left to
// their last writer (the diagonal); that needs lx >= 2 and ly >= 2 (contexts
// are at least 3 x 3).
template <typename S> __device__ __forceinline__ S stored_zero(int i) {
    if constexpr (sizeof(S) == 8) return 0.;
    else return (S)(-(i == 0 ? 4. / 9. : (i <= 4 ? 1. / 9. : 1. / 36.)));
}

template <typename S>
__global__ void k_periodic(Layout L, const S *__restrict__ T, S *__restrict__ Lb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lx = L.lx, ly = L.ly;
    auto at = [&](int i, int x, int y) { return (int64_t)i * L.plane + L.node(x, y); };
    auto ghost = [&](int i, int gx, int gy) -> S {
        const int sx = gx - cx(i), sy = gy - cy(i);
        if (sx >= 1 && sx <= lx && sy >= 1 && sy <= ly) return T[at(i, sx, sy)];
        return stored_zero<S>(i);
    };
    if (t < lx) {
        const int x = t + 1;
        Lb[at(4, x, ly)] = ghost(4, x, 0);
        if (x != lx) Lb[at(7, x, ly)] = ghost(7, x, 0);
        if (x != 1) Lb[at(8, x, ly)] = ghost(8, x, 0);
        Lb[at(2, x, 1)] = ghost(2, x, ly + 1);
        if (x != 1) Lb[at(5, x, 1)] = ghost(5, x, ly + 1);
        if (x != lx) Lb[at(6, x, 1)] = ghost(6, x, ly + 1);
    }
    if (t < ly) {
        const int y = t + 1;
        Lb[at(1, 1, y)] = ghost(1, lx + 1, y);
        if (y != 1) Lb[at(5, 1, y)] = ghost(5, lx + 1, y);
        if (y != ly) Lb[at(8, 1, y)] = ghost(8, lx + 1, y);
        Lb[at(3, lx, y)] = ghost(3, 0, y);
        if (y != 1) Lb[at(6, lx, y)] = ghost(6, 0, y);
        if (y != ly) Lb[at(7, lx, y)] = ghost(7, 0, y);
    }
    if (t == 0) {
        Lb[at(5, 1, 1)] = ghost(5, lx + 1, ly + 1);
        Lb[at(6, lx, 1)] = ghost(6, 0, ly + 1);
        Lb[at(7, lx, ly)] = ghost(7, 0, 0);
        Lb[at(8, 1, ly)] = ghost(8, lx + 1, 0);
    }
}

hipError_t launch_periodic(int storage, const Layout &L, const void *T, void *Lb,
                           hipStream_t st) {
    if (L.lx < 2 || L.ly < 2) return hipErrorInvalidValue;
    const int n = L.lx > L.ly ? L.lx : L.ly;
    const int blocks = (n + 255) / 256;
    if (storage == 0) k_periodic<double><<<blocks, 256, 0, st>>>(L, (const double *)T, (double *)Lb);
    else if (storage == 1) k_periodic<float><<<blocks, 256, 0, st>>>(L, (const float *)T, (float *)Lb);
    else k_periodic<_Float16><<<blocks, 256, 0, st>>>(L, (const _Float16 *)T, (_Float16 *)Lb);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t st) {
    const int64_t n = bytes / 16;
    const int64_t blocks = (n + 255) / 256;
    if (blocks < 1 || blocks > INT32_MAX) return hipErrorInvalidValue;
    k_copy<<<(unsigned)blocks, 256, 0, st>>>((const dbl2 *)src, (dbl2 *)dst, n);
    return hipGetLastError();
}

hipError_t launch_diag_final(const double *partial, int64_t nblocks, double *out, hipStream_t st) {
    k_diag_final<<<1, 256, 0, st>>>(partial, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_node_op_f64(int which, int n, double *f, const double *p0, const double *p1,
                              double omega, hipStream_t st) {
    k_node_op<<<(n + 63) / 64, 64, 0, st>>>(which, n, f, p0, p1, omega);
    return hipGetLastError();
}

}  // namespace lbm
