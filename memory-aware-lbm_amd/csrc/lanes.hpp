// lanes.hpp -- moving values between neighbouring lanes of a wave64.
//
// Lane shifts by one row go through DPP wave shifts (a VALU move, a few
// cycles) instead of ds_bpermute through the LDS crossbar: in the register
// marches the shifted values feed the next level's collision directly, so
// their latency sits on the march's dependency chain (six-step FP64 march
// +2%, four-step FP32 +2%, same bits).  Lane l takes lane l-1 (wave_shr:1,
// lane_up) or l+1 (wave_shl:1, lane_down); the lane without a source keeps
// its own value, as __shfl_up / __shfl_down leave it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
template <int CTRL> __device__ __forceinline__ uint32_t dpp32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ __forceinline__ double dpp64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = dpp32<CTRL>((uint32_t)u), hi = dpp32<CTRL>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, (hi << 32) | lo);
}
__device__ __forceinline__ double lane_up(double v) { return dpp64<DPP_WAVE_SHR1>(v); }
__device__ __forceinline__ double lane_down(double v) { return dpp64<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ float lane_up(float v) {
    return __builtin_bit_cast(float, dpp32<DPP_WAVE_SHR1>(__builtin_bit_cast(uint32_t, v)));
}
__device__ __forceinline__ float lane_down(float v) {
    return __builtin_bit_cast(float, dpp32<DPP_WAVE_SHL1>(__builtin_bit_cast(uint32_t, v)));
}
__device__ __forceinline__ _Float16 lane_up(_Float16 v) {
    const uint32_t u = __builtin_bit_cast(uint16_t, v);
    return __builtin_bit_cast(_Float16, (uint16_t)dpp32<DPP_WAVE_SHR1>(u));
}
__device__ __forceinline__ _Float16 lane_down(_Float16 v) {
    const uint32_t u = __builtin_bit_cast(uint16_t, v);
    return __builtin_bit_cast(_Float16, (uint16_t)dpp32<DPP_WAVE_SHL1>(u));
}

}  // namespace lbm
