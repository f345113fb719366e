// segwalk.h -- the flat segment walk shared by the count engine and the
// per-edge support kernels (gfx950, wave64).
//
// A window holds up to 64 lists ("segments"), one per lane; their lengths are
// scanned into one flat sequence of entries, which a wave consumes in rounds
// of UNROLL chunks of 64 positions.  These helpers find each position's column
// index from the segments' bases and ends held in lane registers, and load the
// entries of a round.
#pragma once
#include "wave.h"

namespace lstc {

// Tables are arrays of 4-slot buckets (16 bytes, aligned); a key's home is a
// bucket (Fibonacci hash: qshift = 34 - log2(slots)).  Linear probing runs
// over buckets and fills a bucket's slots in order, so a bucket whose last slot
// is empty ends every chain through it.
__device__ __forceinline__ uint32_t hquad(uint32_t x, uint32_t qshift) { return (x * 0x9E3779B1u) >> qshift; }

// One round of a window: the entries [B, B + 64*UNROLL) of the window's flat
// sequence (B < total), chunk k (64 entries) loaded into x[k].  Lane i of the
// wave holds segment i: mybase = its column start minus its flat start (mod
// 2^32: column indices are below 2^32 because m < 2^32 - 4096), myend =
// its exclusive flat end (kEmpty for lanes without a segment); ends ascend.
// Positions at or past `total` load col[0] (in bounds, since total > 0 means the
// graph has arcs) and are never probed: every load is unconditional, so its
// result register is never copied and the loads stay in flight until probed.

// The round's first segment: segments that end at or before B lie wholly
// before it.
__device__ __forceinline__ uint32_t first_segment(uint32_t B, uint32_t myend) {
    return (uint32_t)__popcll(ballot(myend <= B));
}
// Whether no segment ends inside the round (an end at B + RB leaves every
// position of the round before it), so the round lies in one segment.
template <int UNROLL>
__device__ __forceinline__ bool one_segment(uint32_t B, uint32_t myend) {
    return ballot(myend > B && myend < B + 64u * UNROLL) == 0;
}

// Column positions are element indices, or with NARROW (every index below
// 2^30) byte offsets, which the loads take as they are.
template <bool NARROW>
__device__ __forceinline__ uint32_t col_pos(uint32_t i) {
    return NARROW ? i << 2 : i;
}

// A full round that lies wholly in one segment: its column index at position
// B + 64k is c0 + 64k (uniform, below m), so only the lane offset is per lane.
template <int UNROLL, bool NARROW>
__device__ __forceinline__ void index_single(uint32_t c0, uint32_t lane, uint32_t (&idx)[UNROLL]) {
    const uint32_t i0 = col_pos<NARROW>(c0 + lane);
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) idx[k] = i0 + col_pos<NARROW>(64u * k);
}

// A round in which segments end: they are the consecutive segments sg, sg + 1,
// ... (ends ascend), so the round walks them with wave-uniform reads of their
// bases and ends.  A chunk inside one segment is addressed from its uniform
// base; in a chunk that a segment end crosses, each lane picks the base of the
// last segment that starts at or before its position.  The walk stops at the
// segment that ends at `total` (the window's last), so j stays below the
// window's segment count.  FULL: the round lies wholly before `total` (every
// round of a window but the last), so no position needs the past-the-end
// handling (index 0 otherwise).
template <int UNROLL, bool FULL, bool NARROW>
__device__ __forceinline__ void index_walk(uint32_t B, uint32_t total, uint32_t mybase, uint32_t myend, uint32_t sg,
                                           uint32_t lane, uint32_t (&idx)[UNROLL]) {
    uint32_t j = sg;
    uint32_t bj = (uint32_t)__builtin_amdgcn_readlane((int)mybase, (int)j);
    uint32_t ej = (uint32_t)__builtin_amdgcn_readlane((int)myend, (int)j);
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) {
        const uint32_t lo = B + 64u * k;
        const uint32_t t = lo + lane;
        uint32_t vb = bj;
        // segments ending inside this chunk (none for most chunks)
        while (ej < lo + 64u && (FULL || ej < total)) {
            const uint32_t e = ej;
            ++j;
            bj = (uint32_t)__builtin_amdgcn_readlane((int)mybase, (int)j);
            ej = (uint32_t)__builtin_amdgcn_readlane((int)myend, (int)j);
            vb = t >= e ? bj : vb;
        }
        idx[k] = FULL || t < total ? col_pos<NARROW>(vb + t) : 0u;
    }
}

// The loads of one round, one per chunk, unconditional (past-the-end positions
// read col[0]).  Every round's loads go through this one code path, so each
// buffer keeps its registers across the branches that chose the indices.
template <int UNROLL, bool NARROW>
__device__ __forceinline__ void load_round(const uint32_t *__restrict__ col, const uint32_t (&idx)[UNROLL],
                                           uint32_t (&x)[UNROLL]) {
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) x[k] = NARROW ? gload_off32(col, idx[k]) : gload(col + idx[k]);
}

}  // namespace lstc
