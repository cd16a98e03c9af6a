// Cross-lane and register helpers shared by the banded DILU kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace fvh {

// Work-group barrier that orders LDS only: global loads issued for later
// steps stay in flight across it (a full __syncthreads waits for them).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1); the lane
// without a source keeps its own `old`.
__device__ inline double dpp_shr1(double v, double old) {  // lane l <- lane l-1; lane 0 <- old
    int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), 0x138, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), 0x138, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double dpp_shl1(double v, double old) {  // lane l <- lane l+1; lane 63 <- old
    int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), 0x130, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), 0x130, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline float dpp_shr1(float v, float old) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ inline float dpp_shl1(float v, float old) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// A register copy the compiler keeps: a value carried to the next step is
// copied out of its prefetch slot before the slot is refilled, so the refill
// lands in the slot's own registers (otherwise the loop rotates the whole
// ring through copies that wait for every load in flight).
__device__ inline double keep(double v) {
    double r;
    asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ inline float keep(float v) {
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

}  // namespace fvh
