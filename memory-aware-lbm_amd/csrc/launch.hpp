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
int64_t diag_partials_needed(const Layout &L);
hipError_t launch_diag_f64(const Problem &p, bool pull, const double *src, double *partial,
                           double *out, hipStream_t st);
hipError_t launch_node_op_f64(int which, int n, double *f, const double *p0, const double *p1,
                              double omega, hipStream_t st);

}  // namespace lbm
