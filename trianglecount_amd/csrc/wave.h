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

// Loads the T at byte offset `byte_off` (32 bits) from a wave-uniform base:
// a scalar base plus one VGPR offset, where a 64-bit address takes two VGPRs
// and a 64-bit add.
template <class T>
__device__ __forceinline__ T gload_off32(const T *base, uint32_t byte_off) {
    return *(const LSTC_GLOBAL_AS T *)((const LSTC_GLOBAL_AS char *)base + byte_off);
}

// Loads through the scalar data cache (constant address space).  Only for
// addresses that are uniform across the wave and memory that nothing writes
// while the kernel runs (written by earlier kernels or copies is fine).
#define LSTC_CONST_AS __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ T sload(const T *p) {
    return *(const LSTC_CONST_AS T *)p;
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
// cnt += this lane's bit of a wave-uniform mask: one vector add whose carry-in
// is the mask (no scalar popcount, no per-lane select).
// The carry-out goes to VCC (clobbered), never to an SGPR the compiler might
// hand to a memory instruction next.
__device__ __forceinline__ void add_lane_bit(uint32_t &cnt, uint64_t mask) {
    asm("v_addc_co_u32_e64 %0, vcc, %1, 0, %2" : "=v"(cnt) : "v"(cnt), "s"(mask) : "vcc");
}
// The inverse of ballot: is this lane's bit set in a wave-uniform mask (the
// mask becomes the execution mask, no vector instruction).
__device__ __forceinline__ bool lane_in(uint64_t mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

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
// Inclusive prefix sum across the wave with DPP (no LDS round trips):
// Hillis-Steele inside each row of 16 lanes (row_shr 1, 2, 4, 8), then lane 15
// of rows 0 and 2 is added to rows 1 and 3 (row_bcast:15, row mask 0b1010) and
// lane 31 to rows 2 and 3 (row_bcast:31, row mask 0b1100).  Lanes whose DPP
// source is out of range or masked receive the `old` operand, 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t /*lane*/) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}
// Sum across the wave (every lane receives it).
__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) { return read_lane(wave_incl_scan(v, 0), 63); }
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
