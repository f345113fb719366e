// lookback.h -- ordered tile prefixes in one pass (decoupled look-back).
//
// Tiles take their index from a ticket, so a tile starts only after every
// earlier one has started.  Each publishes its count as an aggregate, then
// walks back over earlier tiles (64 at a time, one per lane) until it meets
// one that has published its inclusive prefix, sums the aggregates between,
// and publishes its own inclusive prefix.  Status words: flag << 62 | count,
// zero = not published yet (the array is cleared before the launch).  The wait
// is bounded: after kLookBackPolls polls it gives up and sets bit 2 of *flag
// (the caller reports a device error), so no wave can spin forever.
#pragma once
#include "wave.h"

namespace lstc {

static constexpr uint64_t kTsAgg = 1ull << 62, kTsIncl = 2ull << 62, kTsCount = (1ull << 62) - 1;
static constexpr uint32_t kLookBackPolls = 1u << 20;

__device__ __forceinline__ uint64_t status_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Run by one whole wave of the tile: publishes count c of tile `tile` and
// returns the sum of the counts of tiles 0 .. tile-1 (every lane).
__device__ __forceinline__ uint64_t tile_lookback(uint64_t *status, uint64_t tile, uint64_t c, uint32_t *flag,
                                                  uint32_t lane) {
    uint64_t excl = 0;
    if (tile == 0) {
        if (lane == 0) status_store(status, kTsIncl | c);
        return 0;
    }
    if (lane == 0) status_store(status + tile, kTsAgg | c);
    int64_t t = (int64_t)tile - 1;  // lane j examines tile t - j
    uint32_t polls = 0;
    while (true) {
        const int64_t tj = t - (int64_t)lane;
        const uint64_t w = tj >= 0 ? status_load(status + tj) : kTsIncl;  // before tile 0: prefix 0
        const uint64_t mi = ballot((w >> 62) == 2), mz = ballot((w >> 62) == 0);
        const uint64_t first = mi & (0ull - mi);           // nearest tile with its inclusive prefix
        const uint64_t below = first ? first - 1 : ~0ull;  // the tiles between it and this one
        if (mz & below) {                                  // one of them has not published yet
            if (++polls > kLookBackPolls) {
                if (lane == 0) atomicOr(flag, 2u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += wave_sum64(lane_in(below | first) ? (w & kTsCount) : 0ull);
        if (first) break;
        t -= 64;
    }
    if (lane == 0) status_store(status + tile, kTsIncl | (excl + c));
    return excl;
}

}  // namespace lstc
