// host_util.hpp -- host-side helpers shared by the C-ABI translation units:
// error reporting and the node-layout <-> device-plane transfer.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lbm_hip.h"
#include "lbm_common.hpp"

extern thread_local std::string lbm_g_err;


inline int fail(int code, const std::string &msg) {
    lbm_g_err = msg;
    return code;
}

inline int hip_fail(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(e == hipErrorOutOfMemory ? LBM_ERR_NOMEM : LBM_ERR_HIP, m);
}

#define HIPCHK(expr)                                    \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)



// Checks shared by single contexts and batches: sizes, relaxation, side
// operators, outlet and storage.
inline int validate_problem(const lbm_problem *p) {
    if (!p) return fail(LBM_ERR_ARG, "null problem");
    if (p->lx < 3 || p->ly < 3) return fail(LBM_ERR_ARG, "lx and ly must be >= 3");
    if ((int64_t)(p->lx + 2) * (p->ly + 2) > (int64_t)1 << 31)
        return fail(LBM_ERR_ARG, "lattice too large for 32-bit row indexing");
    if (!(p->omega > 0.0 && p->omega <= 2.0))
        return fail(LBM_ERR_ARG, "omega must lie in (0, 2]");
    for (int s = 0; s < 4; ++s) {
        const lbm_side_bc &b = p->side[s];
        if (b.kind != LBM_BC_VELOCITY && b.kind != LBM_BC_PRESSURE)
            return fail(LBM_ERR_ARG, "side kind must be velocity or pressure");
        if (b.family != LBM_BC_REGULARIZED && b.family != LBM_BC_ZOUHE)
            return fail(LBM_ERR_ARG, "side family must be regularized or Zou/He");
    }
    if (p->zgb && p->side[LBM_SIDE_RIGHT].kind != LBM_BC_PRESSURE)
        return fail(LBM_ERR_ARG, "zgb needs a pressure operator on the right side");
    if (p->storage != LBM_STORE_F64 && p->storage != LBM_STORE_F32 && p->storage != LBM_STORE_F16)
        return fail(LBM_ERR_ARG, "unknown storage format");
    return LBM_OK;
}

// host node layout <-> device buffer, interior nodes only.  One population
// at a time: its lx + 2 column segments (seg contiguous elements each, pitch
// apart on the device) go through a pinned staging buffer as one strided copy.
inline int transfer_f64(const lbm::Layout &L, double *host, double *devbuf, bool to_device) {
    const size_t seg = (size_t)L.segment();
    const size_t cols = (size_t)L.lx + 2;
    const size_t seg_bytes = seg * sizeof(double), pitch_bytes = (size_t)L.pitch * sizeof(double);
    double *stage = nullptr;
    HIPCHK(hipHostMalloc((void **)&stage, seg_bytes * cols, hipHostMallocDefault));
    const int ny = L.ly + 2;
    auto at = [&](int x, int y) { return (size_t)x * seg + y + L.yoff; };
    hipError_t e = hipSuccess;
    for (int i = 0; i < 9 && e == hipSuccess; ++i) {
        double *dev = devbuf + (int64_t)i * L.plane;
        // keep the device's ghost ring: read the population first on upload too
        e = hipMemcpy2D(stage, seg_bytes, dev, pitch_bytes, seg_bytes, cols, hipMemcpyDeviceToHost);
        if (e != hipSuccess) break;
        if (to_device) {
            for (int x = 1; x <= L.lx; ++x)
                for (int y = 1; y <= L.ly; ++y)
                    stage[at(x, y)] = host[((size_t)x * ny + y) * 9 + i];
            e = hipMemcpy2D(dev, pitch_bytes, stage, seg_bytes, seg_bytes, cols, hipMemcpyHostToDevice);
        } else {
            for (int x = 1; x <= L.lx; ++x)
                for (int y = 1; y <= L.ly; ++y)
                    host[((size_t)x * ny + y) * 9 + i] = stage[at(x, y)];
        }
    }
    hipError_t e2 = hipHostFree(stage);
    if (e != hipSuccess) return hip_fail(e, "population transfer");
    if (e2 != hipSuccess) return hip_fail(e2, "hipHostFree");
    return LBM_OK;
}
