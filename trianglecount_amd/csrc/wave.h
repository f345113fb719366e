// wave.h -- wave64 helpers for gfx950 (ballot / popcount / lane permutes).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lstc {

// Loads through an explicit global (address space 1) pointer.  Pointers that
// reach a kernel inside a struct are generic to the compiler, which then emits
// FLAT loads; those share the LDS counter, so every LDS wait also drains the
// outstanding global loads.  Global loads keep the two queues independent.
#define LSTC_GLOBAL_AS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gload(const T *p) {
    return *(const LSTC_GLOBAL_AS T *)p;
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// Number of set bits of mask in lanes strictly below this lane.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint32_t read_lane(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
__device__ __forceinline__ uint64_t read_lane64(uint64_t v, uint32_t lane) {
    uint32_t lo = read_lane((uint32_t)v, lane), hi = read_lane((uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}
// Gather v from lane src (any lane, per-lane source) via ds_bpermute.
__device__ __forceinline__ uint32_t gather(uint32_t v, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}
__device__ __forceinline__ uint64_t gather64(uint64_t v, uint32_t src) {
    uint32_t lo = gather((uint32_t)v, src), hi = gather((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
// Inclusive prefix sum across the wave (Hillis-Steele over ds_bpermute).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t y = gather(v, lane >= d ? lane - d : lane);
        if (lane >= d) v += y;
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint64_t y = gather64(v, lane >= d ? lane - d : lane);
        if (lane >= d) v += y;
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (uint32_t d = 32; d >= 1; d >>= 1) v += gather64(v, lane_id() ^ d);
    return v;
}

}  // namespace lstc
