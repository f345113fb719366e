// selftest.hip -- device check of the wave64 primitives the kernels build on.
//
// The counting walk turns DPP prefix sums, lane permutes and ballot ranks into
// load addresses and counts hits by adding ballot masks lane by lane, so a
// wrong lane layout would send loads astray or miscount.  This check
// runs each primitive on fixed patterns, one wave per pattern, and compares the
// lanes with host arithmetic.  Its kernel only writes its own output buffer.
#include <vector>

#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint32_t kPatterns = 64;
static constexpr uint32_t kOutputs = 7;  // value, inclusive scan, sum, reversed gather, odd-rank,
                                         // mask-bit adds, inverse ballot

__device__ __forceinline__ uint32_t pattern_value(uint32_t p, uint32_t lane) {
    switch (p) {
    case 0: return 0u;
    case 1: return 1u;
    case 2: return lane;
    case 3: return lane == 63 ? 7u : 0u;
    case 4: return lane == 0 ? 0xFFFFFFFFu : 0u;
    case 5: return (lane & 15) == 15 ? 1000u : 0u;  // row ends only
    case 6: return 0x80000000u + lane;              // sums wrap mod 2^32
    default: {
        uint32_t x = p * 0x9E3779B9u ^ lane * 0x85EBCA6Bu;
        x ^= x >> 15;
        x *= 0x2C1B3C6Du;
        x ^= x >> 12;
        return (x & 3u) == 0 ? 0u : (x >> (p & 15));  // a quarter zeros, varied magnitudes
    }
    }
}

__global__ __launch_bounds__(64) void k_wave_selftest(uint32_t *__restrict__ out) {
    const uint32_t p = blockIdx.x, lane = lane_id();
    const uint32_t v = pattern_value(p, lane);
    uint32_t *o = out + (size_t)p * kOutputs * 64;
    o[0 * 64 + lane] = v;
    o[1 * 64 + lane] = wave_incl_scan(v, lane);
    o[2 * 64 + lane] = wave_sum32(v);
    o[3 * 64 + lane] = gather(v, 63 - lane);
    o[4 * 64 + lane] = mask_rank(ballot((v & 1u) != 0));
    uint32_t cnt = 3u * lane;
    add_lane_bit(cnt, ballot((v & 2u) != 0));
    add_lane_bit(cnt, ballot((v & 4u) != 0));
    o[5 * 64 + lane] = cnt;
    o[6 * 64 + lane] = lane_in(ballot((v & 1u) != 0) ^ 0x5555555555555555ull) ? 1u : 0u;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_selftest(uint32_t *mismatches) {
    if (!mismatches) return fail(LSTC_ERR_INVALID, "lstc_selftest: NULL output");
    *mismatches = 0;
    const size_t words = (size_t)kPatterns * kOutputs * 64;
    uint32_t *d = nullptr;
    LSTC_HIP(hipMalloc((void **)&d, words * 4));
    hipLaunchKernelGGL(k_wave_selftest, dim3(kPatterns), dim3(64), 0, nullptr, d);
    hipError_t e = hipGetLastError();
    std::vector<uint32_t> h(words);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d, words * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "lstc_selftest");
    uint32_t bad = 0;
    for (uint32_t p = 0; p < kPatterns; ++p) {
        const uint32_t *o = h.data() + (size_t)p * kOutputs * 64;
        uint32_t run = 0, total = 0, odd = 0;
        for (uint32_t l = 0; l < 64; ++l) total += o[l];
        for (uint32_t l = 0; l < 64; ++l) {
            run += o[l];
            bad += o[64 + l] != run;
            bad += o[128 + l] != total;
            bad += o[192 + l] != o[63 - l];
            bad += o[256 + l] != odd;
            odd += o[l] & 1u;
            bad += o[320 + l] != 3u * l + ((o[l] >> 1) & 1u) + ((o[l] >> 2) & 1u);
            bad += o[384 + l] != ((o[l] & 1u) ^ (l & 1u ? 0u : 1u));
        }
    }
    *mismatches = bad;
    return LSTC_OK;
}
