// launch.hpp -- host-side launchers of the device kernels (C++ linkage,
// private to the library; the public surface is include/lbm_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include "lbm_common.hpp"

namespace lbm {

// FP64 (bit-exact) path: kernels_f64.hip
hipError_t launch_collide_f64(const Problem &p, bool pull, const double *src, double *dst,
                              hipStream_t st);
hipError_t launch_stream_f64(const Layout &L, const double *src, double *dst, hipStream_t st);
hipError_t launch_init_f64(const Layout &L, const double *init_ux, double *dst, hipStream_t st);
hipError_t launch_macros_f64(const Layout &L, bool pull, const double *src, double *rho,
                             double *ux, double *uy, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t st);
int64_t diag_partials_needed(const Layout &L);
hipError_t launch_diag_f64(const Problem &p, bool pull, const double *src, double *partial,
                           double *out, hipStream_t st);
hipError_t launch_node_op_f64(int which, int n, double *f, const double *p0, const double *p1,
                              double omega, hipStream_t st);

hipError_t launch_diag_final(const double *partial, int64_t nblocks, double *out, hipStream_t st);

// Reduced-precision storage (FP32 = 1, FP16 = 2): kernels_lowp.hip
hipError_t launch_collide_lp(int storage, const Problem &p, bool pull, const void *src, void *dst,
                             hipStream_t st);
hipError_t launch_stream_lp(int storage, const Layout &L, const void *src, void *dst,
                            hipStream_t st);
hipError_t launch_init_lp(int storage, const Layout &L, const double *init_ux, void *dst,
                          hipStream_t st);
hipError_t launch_macros_lp(int storage, const Layout &L, bool pull, const void *src, double *rho,
                            double *ux, double *uy, hipStream_t st);
hipError_t launch_diag_lp(int storage, const Problem &p, bool pull, const void *src,
                          double *partial, hipStream_t st);
hipError_t launch_fill_zero_f(int storage, int64_t plane, void *dst, hipStream_t st);
hipError_t launch_widen(int storage, int64_t plane, const void *src, double *dst, hipStream_t st);
hipError_t launch_narrow(int storage, int64_t plane, const double *src, void *dst, hipStream_t st);

// Two-step (temporally blocked) kernels: kernels_2step_f64.hip
struct PanelGrid {
    int ty;   // rows owned per panel (<= 254)
    int cx;   // columns per chunk (>= 3)
    int npy, ncx;
};
struct BlockGrid {
    int tx;   // tile columns (>= 3)
    int typ;  // tile rows + 2 (32, 64 or 128)
    int nbx, nby;
};
PanelGrid panel_grid(const Layout &L, int ty, int cx);
BlockGrid block_grid(const Layout &L, int tx, int typ);
size_t block_lds_bytes(const BlockGrid &g);
hipError_t launch_panel_f64(const Problem &p, const PanelGrid &g, const double *P, double *Q,
                            hipStream_t st);
hipError_t launch_block_f64(const Problem &p, const BlockGrid &g, const double *P, double *Q,
                            hipStream_t st);
hipError_t block_kernel_setup();

}  // namespace lbm
