// support.hip -- per-edge triangle support and per-vertex triangle counts.
//
// The reference keeps per-source partials g_sum[u] (the triangles whose
// lowest-ranked vertex is u) before it reduces them
// (src/WithoutPartition/ShareKernelTCV62_update_v1.cu:175-186); local
// clustering and k-truss need the counts per vertex and per edge instead.
// Every triangle u -> w -> x (u -> x, u owned) is found once, as an entry x of
// a streamed list that the hashed list also holds, and adds one to the support
// of its three arcs and to the count of its three vertices.
//
// A work item hashes one row r (in tiles of at most half its table's slots)
// into an LDS table of 4-slot buckets -- a wave's slice of the workgroup's
// table, or for rows longer than a slice holds the whole table shared by the
// workgroup's waves -- and streams candidate lists through it with the count
// engine's flat segment walk (segwalk.h): 64 lists per window, rounds of
// kSupUnroll chunks of 64 entries, a round inside one list addressed from a
// scalar base, the next round's loads issued before this one is probed.  A hit
// (entry x of the list of candidate c) is attributed three ways:
//   arc r -> x : the hit slot's counter (one LDS add; the slot word also holds
//                the key's index in the tile, so the counters go out to the
//                supports with one atomic per slot and tile);
//   arc of c   : hits per list come from a running count of the window's hits
//                read at each list's end (arc r -> w for a neighbour w of r;
//                arc u -> r for an in-list source u, located by one binary
//                search per list with hits), one atomic per list;
//   the entry  : one global atomic on the support of the streamed arc.
// Vertex counts take the same three totals (x per slot, c per list, r per item).
#include "lstc_internal.h"
#include "segwalk.h"
#include "wave.h"

namespace lstc {

// A workgroup of WAVES waves (4 for the split rule, 8 for the plain walk: the
// faster of the two for each at RMAT s23) holds one table of 2^kSupGroupLog2
// slots, or one slice of it per wave.
static constexpr uint32_t kSupGroupLog2 = 12;                             // slots of a workgroup's table
static constexpr uint32_t kSupGroupTile = (1u << kSupGroupLog2) / 2;      // keys per tile: load <= 1/2
template <int WAVES>
constexpr uint32_t sup_wave_log2() { return kSupGroupLog2 - (WAVES == 8 ? 3 : 2); }  // slots of a wave's table
template <int WAVES>
constexpr uint32_t sup_wave_tile() { return (1u << sup_wave_log2<WAVES>()) / 2; }
static constexpr int kSupSplitWaves = 4, kSupPlainWaves = 8;
static constexpr int kSupUnroll = 4;                                       // chunks of 64 entries per round
static constexpr uint32_t kSupEmpty = 0xFFFFFFFFu;     // empty slot; also the end of a lane without a list
static constexpr uint32_t kSupNoId = 0xFFFFFFFEu;      // past-the-end positions: never a key (ids < 2^32 - 2)
static constexpr uint32_t kIdxShift = 21;              // slot word: tile index << 21 | triangles found
static constexpr uint32_t kHitMask = (1u << kIdxShift) - 1u;
static constexpr uint32_t kSupPass = 1u << 20;         // candidates per pass: a slot gains <= 1 per list
static_assert(kSupGroupTile <= (1u << (32 - kIdxShift)) && kSupPass <= kHitMask, "slot word fields");

// A table: keys in 4-slot buckets, one slot word per slot (arc r -> key), the
// build's overflow bits; a wave's slice of the workgroup's arrays or all of them.
struct SupTab {
    uint32_t *tab, *cnt, *ovf;
};
// A window's non-empty lists: j-th list's column start minus its flat start,
// its flat end (exclusive), its candidate's lane in the window.
struct SupWin {
    uint32_t *base, *end;
    uint8_t *cand;
};

// log2 of the table slots for T keys: at least twice as many slots, 16 at least.
__device__ __forceinline__ uint32_t sup_tlog2(uint32_t T) {
    uint32_t l = 4;
    while ((1u << l) < 2 * T) ++l;
    return l;
}

// Orders the LDS accesses of the NT threads that share a table: a workgroup
// barrier, or for one wave (whose LDS accesses issue in order) a compiler
// barrier only.
template <int NT>
__device__ __forceinline__ void sup_sync() {
    if (NT > 64) __syncthreads();
    else __builtin_amdgcn_wave_barrier();
}

// Builds the table of the tile col[kb, kb + T) with NT threads: a key takes the
// first free slot of its home bucket or, when that is full, of the buckets
// after it (a slot is tried once the ones before it are taken; at most half the
// slots are filled, so one lap finds a free one), and its slot word gets its
// tile index.  A home that overflowed is then marked by the order of its first
// two slots (such a bucket is full: slot 0 > slot 1; any other: slot 0 < slot
// 1, or slot 1 empty), so one read of the home tells a probe whether to go on.
template <int NT>
__device__ __forceinline__ void sup_build(const SupTab &Q, const uint32_t *__restrict__ col, uint64_t kb, uint32_t T,
                                          uint32_t tlog2, uint32_t tid) {
    const uint32_t ns = 1u << tlog2, qmask = (ns >> 2) - 1u, qshift = 34 - tlog2;
    for (uint32_t i = tid; i < ns; i += NT) {
        Q.tab[i] = kSupEmpty;
        Q.cnt[i] = 0;
    }
    for (uint32_t i = tid; i < (ns + 127) >> 7; i += NT) Q.ovf[i] = 0;
    sup_sync<NT>();
    for (uint32_t i = tid; i < T; i += NT) {
        const uint32_t key = gload(col + kb + i);
        const uint32_t home = hquad(key, qshift);
        uint32_t h = home;
        for (uint32_t step = 0; step <= qmask; ++step) {
            uint32_t *q = Q.tab + 4 * h;
            int sl = -1;
            if (atomicCAS(q, kSupEmpty, key) == kSupEmpty) sl = 0;
            else if (atomicCAS(q + 1, kSupEmpty, key) == kSupEmpty) sl = 1;
            else if (atomicCAS(q + 2, kSupEmpty, key) == kSupEmpty) sl = 2;
            else if (atomicCAS(q + 3, kSupEmpty, key) == kSupEmpty) sl = 3;
            if (sl >= 0) {
                Q.cnt[4 * h + sl] = i << kIdxShift;
                break;
            }
            h = (h + 1) & qmask;
        }
        if (h != home) atomicOr(Q.ovf + (home >> 5), 1u << (home & 31u));
    }
    sup_sync<NT>();
    for (uint32_t b = tid; b <= qmask; b += NT) {
        const uint32_t k0 = Q.tab[4 * b], k1 = Q.tab[4 * b + 1];
        const bool over = (Q.ovf[b >> 5] >> (b & 31u)) & 1u;
        if (over ? k0 < k1 : k0 > k1) {
            const uint32_t w0 = Q.cnt[4 * b], w1 = Q.cnt[4 * b + 1];
            Q.tab[4 * b] = k1;
            Q.tab[4 * b + 1] = k0;
            Q.cnt[4 * b] = w1;
            Q.cnt[4 * b + 1] = w0;
        }
    }
    sup_sync<NT>();
}

// Sends the slot counters out with NT threads (arcs tile_arc + tile index,
// vertices = the keys) and clears them; the tile indexes stay.
template <int NT>
__device__ __forceinline__ void sup_flush(const SupTab &Q, uint64_t tile_arc, uint32_t tlog2,
                                          uint32_t *__restrict__ support, unsigned long long *__restrict__ per_vertex,
                                          uint32_t tid) {
    sup_sync<NT>();
    for (uint32_t i = tid; i < (1u << tlog2); i += NT) {
        const uint32_t w = Q.cnt[i];
        const uint32_t h = w & kHitMask;
        if (h == 0) continue;
        if (support) atomicAdd(support + tile_arc + (w >> kIdxShift), h);
        if (per_vertex) atomicAdd(per_vertex + Q.tab[i], (unsigned long long)h);
        Q.cnt[i] = w & ~kHitMask;
    }
    sup_sync<NT>();
}

// Adds one to the support of the streamed arc at column position i (a byte
// offset with NARROW).
template <bool NARROW>
__device__ __forceinline__ void sup_entry_add(uint32_t *__restrict__ support, uint32_t i) {
    LSTC_GLOBAL_AS uint32_t *p = NARROW ? (LSTC_GLOBAL_AS uint32_t *)((LSTC_GLOBAL_AS char *)support + i)
                                        : (LSTC_GLOBAL_AS uint32_t *)support + i;
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Probes the chunks of one round (positions B + 64k + lane, entries x[k] at
// column positions idx[k]) and attributes the hits: slot counters and the
// streamed arcs at once; for the lists, `cum` counts the window's hits so far
// and a list that ends in this round (ends != 0: some list does) records in its
// lane the count at its end.  FULL: the round lies wholly before `total`.
template <int U, bool FULL, bool NARROW>
__device__ __forceinline__ void sup_probe(const SupTab &Q, const uint32_t (&x)[U], const uint32_t (&idx)[U],
                                          uint32_t B, uint32_t total, uint64_t ends, uint32_t myend, uint32_t qmask,
                                          uint32_t qshift, uint32_t *__restrict__ support, uint32_t lane,
                                          uint32_t &cum, uint32_t &cend) {
    const uint4 *bk = (const uint4 *)Q.tab;
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const uint32_t lo = B + 64u * k;
        const bool valid = FULL || lo + lane < total;
        const uint32_t xv = valid ? x[k] : kSupNoId;
        const uint32_t p = hquad(xv, qshift);
        const uint4 v = bk[p];
        uint64_t H = ballot(v.x == xv) | ballot(v.y == xv) | ballot(v.z == xv) | ballot(v.w == xv);
        uint32_t slot = 4 * p + (v.y == xv ? 1u : v.z == xv ? 2u : v.w == xv ? 3u : 0u);
        // no hit and the home overflowed: the chain goes on (within one lap)
        uint64_t going = ~(H | ballot(v.x <= v.y));
        if (!FULL) going &= ballot(valid);
        if (going != 0) {
            uint32_t q = p;
            for (uint32_t step = 0; step < qmask && going != 0; ++step) {
                q = (q + 1) & qmask;
                const uint4 w = bk[q];
                const uint64_t G = (ballot(w.x == xv) | ballot(w.y == xv) | ballot(w.z == xv) | ballot(w.w == xv)) &
                                   going;
                if (lane_in(G)) slot = 4 * q + (w.y == xv ? 1u : w.z == xv ? 2u : w.w == xv ? 3u : 0u);
                H |= G;
                going &= ~(G | ballot(w.w == kSupEmpty));
            }
        }
        if (lane_in(H)) {
            atomicAdd(Q.cnt + slot, 1u);
            if (support) sup_entry_add<NARROW>(support, idx[k]);
        }
        if (ends != 0) {  // lists ending at lo + d, 1 <= d <= 64: the hits before their end
            const uint32_t d = myend - lo;
            const uint32_t at = cum + (uint32_t)__popcll(H << ((64u - d) & 63u));  // kept when 1 <= d <= 64
            cend = d - 1u < 64u ? at : cend;
        }
        cum += (uint32_t)__popcll(H);
    }
}

// Column positions of round B (index_walk / index_single); returns the lanes
// whose list ends inside the round or at its last position.
template <int U, bool NARROW>
__device__ __forceinline__ uint64_t sup_index(uint32_t B, uint32_t total, uint32_t mybase, uint32_t myend,
                                              uint32_t lane, uint32_t (&idx)[U]) {
    constexpr uint32_t RB = 64u * U;
    const uint32_t sg = first_segment(B, myend);
    if (B + RB > total)
        index_walk<U, false, NARROW>(B, total, mybase, myend, sg, lane, idx);
    else if (one_segment<U>(B, myend))
        index_single<U, NARROW>(read_lane(mybase, sg) + B, lane, idx);
    else
        index_walk<U, true, NARROW>(B, total, mybase, myend, sg, lane, idx);
    return ballot(myend > B && myend <= B + RB);
}

// Streams one window: each lane offers the list col[s, s + len) (len 0: none)
// of its candidate.  Returns in lane j the triangles found in the window's j-th
// non-empty list (whose candidate's lane is W.cand[j]); nseg = their number.
template <int U, bool NARROW>
__device__ __forceinline__ uint32_t sup_window(const SupTab &Q, const SupWin &W, const uint32_t *__restrict__ col,
                                               uint32_t *__restrict__ support, uint32_t s, uint32_t len,
                                               uint32_t qmask, uint32_t qshift, uint32_t lane, uint32_t &nseg) {
    constexpr uint32_t RB = 64u * U;
    const uint32_t incl = wave_incl_scan(len, lane);
    const uint32_t total = read_lane(incl, 63);  // the lists are disjoint ranges of col: total <= m
    const uint64_t nzm = ballot(len != 0);
    nseg = (uint32_t)__popcll(nzm);
    if (total == 0) return 0;
    if (len != 0) {
        const uint32_t k = mask_rank(nzm);
        W.base[k] = s - (incl - len);
        W.end[k] = incl;
        W.cand[k] = (uint8_t)lane;
    }
    const uint32_t mybase = W.base[lane];
    const uint32_t myend = lane < nseg ? W.end[lane] : kSupEmpty;
    uint32_t cum = 0, cend = 0;
    uint32_t ia[U], ib[U], xa[U], xb[U];
    uint64_t ea = sup_index<U, NARROW>(0, total, mybase, myend, lane, ia), eb;
    load_round<U, NARROW>(col, ia, xa);
    uint32_t B = 0;
    for (; B + 3 * RB <= total; B += 2 * RB) {  // rounds B + RB and B + 2 RB are full
        eb = sup_index<U, NARROW>(B + RB, total, mybase, myend, lane, ib);
        load_round<U, NARROW>(col, ib, xb);
        sup_probe<U, true, NARROW>(Q, xa, ia, B, total, ea, myend, qmask, qshift, support, lane, cum, cend);
        ea = sup_index<U, NARROW>(B + 2 * RB, total, mybase, myend, lane, ia);
        load_round<U, NARROW>(col, ia, xa);
        sup_probe<U, true, NARROW>(Q, xb, ib, B + RB, total, eb, myend, qmask, qshift, support, lane, cum, cend);
    }
    // round B is in xa; rounds B + RB and B + 2 RB may follow, the last one partial
    if (B + RB >= total) {
        sup_probe<U, false, NARROW>(Q, xa, ia, B, total, ea, myend, qmask, qshift, support, lane, cum, cend);
    } else {
        eb = sup_index<U, NARROW>(B + RB, total, mybase, myend, lane, ib);
        load_round<U, NARROW>(col, ib, xb);
        sup_probe<U, true, NARROW>(Q, xa, ia, B, total, ea, myend, qmask, qshift, support, lane, cum, cend);
        if (B + 2 * RB < total) {
            ea = sup_index<U, NARROW>(B + 2 * RB, total, mybase, myend, lane, ia);
            load_round<U, NARROW>(col, ia, xa);
            sup_probe<U, true, NARROW>(Q, xb, ib, B + RB, total, eb, myend, qmask, qshift, support, lane, cum, cend);
            sup_probe<U, false, NARROW>(Q, xa, ia, B + 2 * RB, total, ea, myend, qmask, qshift, support, lane, cum,
                                        cend);
        } else {
            sup_probe<U, false, NARROW>(Q, xb, ib, B + RB, total, eb, myend, qmask, qshift, support, lane, cum, cend);
        }
    }
    // list j's triangles: the count at its end minus the count at the end of list j - 1
    const uint32_t prev = gather(cend, lane - 1u);
    return lane < nseg ? cend - (lane ? prev : 0u) : 0u;
}

// Position of key r in the sorted list col[b, e) (r is there).
__device__ __forceinline__ uint64_t sup_position(const uint32_t *__restrict__ col, uint64_t b, uint64_t e, uint32_t r) {
    while (e - b > 1) {
        const uint64_t mid = (b + e) >> 1;
        if (col[mid] <= r) b = mid; else e = mid;
    }
    return b;
}

// One work item at row r (N+(r) = col[rb, rb + K)), run by NT threads (one
// wave, or the workgroup sharing one table, whose waves claim windows from the
// counter *next).  Candidates [0, Ka): kind A, the c-th neighbour w of r,
// N+(w) streamed (only when d+(w) < below, if below != 0); candidates
// [Ka, Ka + Kb): kind B, the in-list source u = src[c - Ka], N+(u) streamed.
// N+(r) is hashed in tiles of TILE keys (a wave tile, or kSupGroupTile); the
// candidates go in passes of at most kSupPass, each followed by a flush of the
// slot counters.
template <int U, bool NARROW, int NT, uint32_t TILE>
__device__ __forceinline__ void sup_item(const SupTab &Q, const SupWin &W, const uint64_t *__restrict__ off,
                                         const uint32_t *__restrict__ col, const uint32_t *__restrict__ src,
                                         uint32_t n_rows, uint32_t r, uint64_t rb, uint32_t K, uint32_t Ka,
                                         uint32_t Kb, uint32_t below, uint32_t *__restrict__ support,
                                         unsigned long long *__restrict__ per_vertex, uint32_t *next, uint32_t tid) {
    const uint32_t lane = tid & 63;
    const uint32_t C = Ka + Kb;
    unsigned long long row_tri = 0;
    for (uint32_t t0 = 0; t0 < K; t0 += TILE) {
        const uint32_t T = min(K - t0, TILE);
        const uint32_t tl = sup_tlog2(T), qmask = (1u << (tl - 2)) - 1u, qshift = 34 - tl;
        sup_build<NT>(Q, col, rb + t0, T, tl, tid);
        const bool tiled = K > TILE;
        const uint32_t lo = gload(col + rb + t0), hi = gload(col + rb + t0 + T - 1);
        for (uint32_t p0 = 0; p0 < C; p0 += kSupPass) {
            const uint32_t p1 = min(C, p0 + kSupPass);
            if (NT > 64) {
                if (tid == 0) *next = p0;
                __syncthreads();
            }
            for (uint32_t c0 = p0;; c0 += 64) {
                if (NT > 64) {  // the workgroup's next window
                    uint32_t t = 0;
                    if (lane == 0) t = atomicAdd(next, 64u);
                    c0 = read_lane(t, 0);
                }
                if (c0 >= p1) break;
                const uint32_t c = c0 + lane;
                uint32_t s = 0, len = 0;
                if (c < min(Ka, p1)) {
                    const uint32_t w = gload(col + rb + c);
                    if (w < n_rows) {
                        const uint64_t s0 = gload(off + w);
                        s = (uint32_t)s0;
                        len = (uint32_t)(gload(off + w + 1) - s0);
                        if (below && len >= below) len = 0;  // that arc streams N+(r), at w
                    }
                } else if (c < p1) {
                    const uint32_t u = gload(src + (c - Ka));
                    if (u < n_rows) {
                        const uint64_t s0 = gload(off + u);
                        s = (uint32_t)s0;
                        len = (uint32_t)(gload(off + u + 1) - s0);
                    }
                }
                // a list that cannot meet this tile is skipped (only when N+(r) is tiled)
                if (tiled && len && (gload(col + s + len - 1) < lo || gload(col + s) > hi)) len = 0;
                uint32_t nseg;
                const uint32_t h = sup_window<U, NARROW>(Q, W, col, support, s, len, qmask, qshift, lane, nseg);
                if (lane < nseg && h != 0) {
                    const uint32_t cc = c0 + W.cand[lane];
                    if (cc < Ka) {  // arc r -> w
                        const uint64_t a = rb + cc;
                        if (support) atomicAdd(support + a, h);
                        if (per_vertex) atomicAdd(per_vertex + gload(col + a), (unsigned long long)h);
                    } else {  // arc u -> r
                        const uint32_t u = gload(src + (cc - Ka));
                        if (support) atomicAdd(support + sup_position(col, gload(off + u), gload(off + u + 1), r), h);
                        if (per_vertex) atomicAdd(per_vertex + u, (unsigned long long)h);
                    }
                    row_tri += h;
                }
            }
            sup_flush<NT>(Q, rb + t0, tl, support, per_vertex, tid);
        }
    }
    const unsigned long long tot = wave_sum64(row_tri);
    if (per_vertex && lane == 0 && tot) atomicAdd(per_vertex + r, tot);
}

// LDS of a support workgroup: the wave tables, or one table of all of them;
// one window per wave.  Under 40 KB: four workgroups per CU.
static constexpr uint32_t kSupScanRows = 256;  // rows checked per step for workgroup rows (plain walk)
template <int WAVES>
struct SupShared {
    uint32_t tab[1u << kSupGroupLog2];
    uint32_t cnt[1u << kSupGroupLog2];
    uint32_t ovf[(1u << kSupGroupLog2) / 128];
    uint32_t base[WAVES][64], end[WAVES][64];
    uint8_t cand[WAVES][64];
    uint32_t rows[kSupScanRows];
    uint32_t nrows, ticket, next;
};

template <int WAVES>
__device__ __forceinline__ SupTab sup_group_tab(SupShared<WAVES> &L) { return SupTab{L.tab, L.cnt, L.ovf}; }
template <int WAVES>
__device__ __forceinline__ SupTab sup_wave_tab(SupShared<WAVES> &L, uint32_t wv) {
    constexpr uint32_t l2 = sup_wave_log2<WAVES>();
    return SupTab{L.tab + (wv << l2), L.cnt + (wv << l2), L.ovf + (wv << l2) / 128};
}
template <int WAVES>
__device__ __forceinline__ SupWin sup_win(SupShared<WAVES> &L, uint32_t wv) {
    return SupWin{L.base[wv], L.end[wv], L.cand[wv]};
}

// Plain walk: every neighbour kind A.  Rows with d+ above a wave tile are found
// kSupScanRows at a time and run by the whole workgroup; then each wave takes
// the rows with 2 <= d+ <= the wave tile.
template <bool NARROW, int WAVES>
__global__ __launch_bounds__(WAVES * 64, WAVES == 8 ? 8 : 4) void k_support(const uint64_t *__restrict__ off,
                                                                           const uint32_t *__restrict__ col,
                                                                           uint32_t n_rows, uint32_t n_owned,
                                                                           uint32_t *__restrict__ support,
                                                                           unsigned long long *__restrict__ per_vertex) {
    constexpr uint32_t WT = sup_wave_tile<WAVES>();
    __shared__ SupShared<WAVES> L;
    const uint32_t tid = threadIdx.x, wv = tid >> 6;
    const SupWin W = sup_win(L, wv);
    for (uint64_t b0 = (uint64_t)blockIdx.x * kSupScanRows; b0 < n_owned; b0 += (uint64_t)gridDim.x * kSupScanRows) {
        if (tid == 0) L.nrows = 0;
        __syncthreads();
        const uint64_t u = b0 + tid;
        if (tid < kSupScanRows && u < n_owned && off[u + 1] - off[u] > WT)
            L.rows[atomicAdd(&L.nrows, 1u)] = (uint32_t)u;
        __syncthreads();
        const uint32_t nb = L.nrows;
        for (uint32_t i = 0; i < nb; ++i) {
            const uint32_t r = L.rows[i];
            const uint64_t rb = off[r];
            const uint32_t K = (uint32_t)(off[r + 1] - rb);
            sup_item<kSupUnroll, NARROW, WAVES * 64, kSupGroupTile>(sup_group_tab(L), W, off, col, nullptr, n_rows, r,
                                                                    rb, K, K, 0, 0, support, per_vertex, &L.next, tid);
        }
        __syncthreads();
    }
    const SupTab Q = sup_wave_tab(L, wv);
    const uint64_t nw = (uint64_t)gridDim.x * WAVES;
    for (uint64_t u = (uint64_t)blockIdx.x * WAVES + wv; u < n_owned; u += nw) {
        const uint64_t rb = off[u];
        const uint32_t K = (uint32_t)(off[u + 1] - rb);
        if (K < 2 || K > WT) continue;
        sup_item<kSupUnroll, NARROW, 64, WT>(Q, W, off, col, nullptr, n_rows, (uint32_t)u, rb, K, K, 0, 0, support,
                                             per_vertex, nullptr, tid & 63);
    }
}

// ---------------------------------------------------------------- split rule
// With target-side in-lists the same outputs come from the split rule of the
// count (count_engine.hip): arc u -> w is handled at u (kind A: N+(u) in the
// table, N+(w) streamed) when d+(w) < d+(u), else at w (kind B: N+(w) in the
// table, N+(u) streamed for u in w's in-list), so every arc streams its shorter
// list (at RMAT s23 2.9e10 entries instead of 7.3e10).  Work items: chunk 0 of
// a row holds its kind A side, chunk c its in-list sources [512c, 512c + 512),
// so a hub's long in-list is spread over waves.  Items of rows with d+ above a
// wave tile ("group" items) are listed first and run by whole workgroups, then
// the rest by single waves; both take items from a ticket.
static constexpr uint32_t kSupChunk = 512;

// Items per row, group rows and wave rows counted apart.
__global__ void k_sup_items(const uint64_t *__restrict__ off, const uint64_t *__restrict__ in_off, uint32_t n_rows,
                            uint32_t n_owned, uint32_t wave_tile, uint32_t *__restrict__ items_g,
                            uint32_t *__restrict__ items_w) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t K = off[r + 1] - off[r], kb = in_off[r + 1] - in_off[r];
        const uint32_t nb = (uint32_t)((kb + kSupChunk - 1) / kSupChunk);
        const uint32_t a = (r < n_owned && K >= 2) ? 1u : 0u;
        const uint32_t it = K == 0 ? 0u : (nb > a ? nb : a);
        items_g[r] = K > wave_tile ? it : 0u;
        items_w[r] = K > wave_tile ? 0u : it;
    }
}

// item_row[t]: group items at [0, n_g), wave items at n_g + [0, n_w).
__global__ void k_sup_item_rows(const uint32_t *__restrict__ items_g, const uint32_t *__restrict__ items_w,
                                const uint64_t *__restrict__ off_g, const uint64_t *__restrict__ off_w,
                                uint32_t n_rows, uint32_t *__restrict__ item_row) {
    const uint64_t n_g = off_g[n_rows];
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows; r += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t c = 0; c < items_g[r]; ++c) item_row[off_g[r] + c] = (uint32_t)r;
        for (uint32_t c = 0; c < items_w[r]; ++c) item_row[n_g + off_w[r] + c] = (uint32_t)r;
    }
}

template <bool NARROW, int NT, uint32_t TILE>
__device__ __forceinline__ void sup_split_item(const SupTab &Q, const SupWin &W, const uint64_t *__restrict__ off,
                                               const uint32_t *__restrict__ col, const uint64_t *__restrict__ in_off,
                                               const uint32_t *__restrict__ in_col, uint32_t n_rows, uint32_t n_owned,
                                               uint32_t r, uint32_t chunk, uint32_t *__restrict__ support,
                                               unsigned long long *__restrict__ per_vertex, uint32_t *next,
                                               uint32_t tid) {
    const uint64_t rb = off[r];
    const uint32_t K = (uint32_t)(off[r + 1] - rb);
    const uint64_t ib = in_off[r], kb = in_off[r + 1] - ib;
    const uint32_t Ka = (chunk == 0 && r < n_owned && K >= 2) ? K : 0u;
    const uint64_t b0 = (uint64_t)chunk * kSupChunk;
    const uint32_t Kb = b0 < kb ? (uint32_t)min(kb - b0, (uint64_t)kSupChunk) : 0u;
    if (Ka + Kb == 0) return;
    sup_item<kSupUnroll, NARROW, NT, TILE>(Q, W, off, col, in_col + ib + b0, n_rows, r, rb, K, Ka, Kb, K, support,
                                           per_vertex, next, tid);
}

template <bool NARROW, int WAVES>
__global__ __launch_bounds__(WAVES * 64, WAVES == 8 ? 8 : 4) void k_support_split(
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ col, const uint64_t *__restrict__ in_off,
    const uint32_t *__restrict__ in_col, uint32_t n_rows, uint32_t n_owned, const uint64_t *__restrict__ off_g,
    const uint64_t *__restrict__ off_w, const uint32_t *__restrict__ item_row, uint32_t *__restrict__ ticket,
    uint32_t *__restrict__ support, unsigned long long *__restrict__ per_vertex) {
    __shared__ SupShared<WAVES> L;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const SupWin W = sup_win(L, wv);
    const uint32_t n_g = (uint32_t)off_g[n_rows], n_w = (uint32_t)off_w[n_rows];
    for (;;) {  // group items (every thread of the workgroup reaches the same break)
        if (tid == 0) L.ticket = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t t = L.ticket;
        __syncthreads();
        if (t >= n_g) break;
        const uint32_t r = item_row[t];
        sup_split_item<NARROW, WAVES * 64, kSupGroupTile>(sup_group_tab(L), W, off, col, in_off, in_col, n_rows,
                                                          n_owned, r, t - (uint32_t)off_g[r], support, per_vertex,
                                                          &L.next, tid);
    }
    const SupTab Q = sup_wave_tab(L, wv);
    for (;;) {  // wave items (every wave reaches this break once the items are taken)
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(ticket + 1, 1u);
        t = read_lane(t, 0);
        if (t >= n_w) break;
        const uint32_t r = item_row[n_g + t];
        sup_split_item<NARROW, 64, sup_wave_tile<WAVES>()>(Q, W, off, col, in_off, in_col, n_rows, n_owned, r,
                                                           t - (uint32_t)off_w[r], support, per_vertex, nullptr, lane);
    }
}

struct SupLayout {
    uint32_t *items_g, *items_w, *item_row, *ticket;
    uint64_t *off_g, *off_w;
    void *scan_ws;
    size_t scan_bytes, bytes;
};

static SupLayout sup_layout(void *ws, size_t ws_bytes, uint32_t n_rows, uint64_t m) {
    Carver c(ws, ws_bytes);
    SupLayout L;
    L.items_g = c.take<uint32_t>((size_t)n_rows + 1);
    L.items_w = c.take<uint32_t>((size_t)n_rows + 1);
    L.off_g = c.take<uint64_t>((size_t)n_rows + 1);
    L.off_w = c.take<uint64_t>((size_t)n_rows + 1);
    L.item_row = c.take<uint32_t>(2 * (size_t)n_rows + m / kSupChunk + 1);
    L.ticket = c.take<uint32_t>(4);
    L.scan_bytes = scan_ws_bytes((uint64_t)n_rows + 1);
    L.scan_ws = c.take<char>(L.scan_bytes);
    L.bytes = c.need();
    return L;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_triangle_support(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows,
                                             uint32_t n_owned, uint64_t m, uint32_t *d_support,
                                             uint64_t *d_per_vertex, lstc_stream s) {
    TraceRange range_("lstc_triangle_support");
    hipStream_t st = (hipStream_t)s;
    if (n_owned > n_rows) return fail(LSTC_ERR_INVALID, "lstc_triangle_support: n_owned %u > n_rows %u", n_owned, n_rows);
    if (n_rows == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_triangle_support: n_rows must be < 2^32-1");
    if (n_rows && !d_off) return fail(LSTC_ERR_INVALID, "lstc_triangle_support: NULL offsets");
    if (m && !d_col) return fail(LSTC_ERR_INVALID, "lstc_triangle_support: NULL columns");
    if (!d_support && !d_per_vertex) return fail(LSTC_ERR_INVALID, "lstc_triangle_support: no output");
    if (m >= 0xFFFFF000ull)  // 32-bit column positions in the walk
        return fail(LSTC_ERR_INVALID, "lstc_triangle_support: m = %llu arcs exceeds 2^32-4096", (unsigned long long)m);
    if (d_support && m) LSTC_TRY(fill_u32(d_support, 0, m, st));
    if (d_per_vertex && n_rows) LSTC_TRY(fill_u64(d_per_vertex, 0, n_rows, st));
    if (n_owned == 0 || m == 0) return LSTC_OK;
    const uint32_t grid = (uint32_t)hmin<uint64_t>(((uint64_t)n_owned + kSupPlainWaves - 1) / kSupPlainWaves,
                                                   (uint64_t)device_cus() * 4);  // four workgroups per CU (LDS)
    if (m < (1ull << 30))  // column positions as 32-bit byte offsets
        hipLaunchKernelGGL((k_support<true, kSupPlainWaves>), dim3(grid), dim3(kSupPlainWaves * 64), 0, st, d_off, d_col,
                           n_rows, n_owned,
                           d_support, (unsigned long long *)d_per_vertex);
    else
        hipLaunchKernelGGL((k_support<false, kSupPlainWaves>), dim3(grid), dim3(kSupPlainWaves * 64), 0, st, d_off,
                           d_col, n_rows, n_owned,
                           d_support, (unsigned long long *)d_per_vertex);
    LSTC_LAUNCHED("k_support");
    return LSTC_OK;
}

extern "C" lstc_status lstc_triangle_support_split_workspace(uint32_t n_rows, uint64_t m, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split_workspace: NULL bytes");
    *bytes = sup_layout(nullptr, 0, n_rows, m).bytes;
    return LSTC_OK;
}

extern "C" lstc_status lstc_triangle_support_split(const uint64_t *d_off, const uint32_t *d_col,
                                                   const uint64_t *d_in_off, const uint32_t *d_in_col, uint32_t n_rows,
                                                   uint32_t n_owned, uint64_t m, uint32_t *d_support,
                                                   uint64_t *d_per_vertex, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_triangle_support_split");
    hipStream_t st = (hipStream_t)s;
    if (n_owned > n_rows)
        return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: n_owned %u > n_rows %u", n_owned, n_rows);
    if (n_rows == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: n_rows must be < 2^32-1");
    if (n_rows && (!d_off || !d_in_off)) return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: NULL offsets");
    if (m && (!d_col || !d_in_col)) return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: NULL columns");
    if (!d_support && !d_per_vertex) return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: no output");
    if (m >= 0xFFFFF000ull)  // 32-bit column positions in the walk
        return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: m = %llu arcs exceeds 2^32-4096",
                    (unsigned long long)m);
    if (2 * (uint64_t)n_rows + m / kSupChunk + 1 >= 0xFFFFFFFFull)
        return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: too many work items");
    const SupLayout L = sup_layout(d_ws, ws_bytes, n_rows, m);
    if (!d_ws || ws_bytes < L.bytes)
        return fail(LSTC_ERR_WORKSPACE, "lstc_triangle_support_split: workspace %zu < %zu", ws_bytes, L.bytes);
    if (d_support && m) LSTC_TRY(fill_u32(d_support, 0, m, st));
    if (d_per_vertex && n_rows) LSTC_TRY(fill_u64(d_per_vertex, 0, n_rows, st));
    if (n_rows == 0 || m == 0) return LSTC_OK;
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n_rows + 255) / 256, (uint64_t)device_cus() * 16);
    hipLaunchKernelGGL(k_sup_items, dim3(grid_n), dim3(256), 0, st, d_off, d_in_off, n_rows, n_owned,
                       sup_wave_tile<kSupSplitWaves>(), L.items_g, L.items_w);
    LSTC_LAUNCHED("k_sup_items");
    LSTC_TRY(scan_exclusive_u32_to_u64(L.items_g, L.off_g, n_rows, L.off_g + n_rows, L.scan_ws, L.scan_bytes, st));
    LSTC_TRY(scan_exclusive_u32_to_u64(L.items_w, L.off_w, n_rows, L.off_w + n_rows, L.scan_ws, L.scan_bytes, st));
    hipLaunchKernelGGL(k_sup_item_rows, dim3(grid_n), dim3(256), 0, st, L.items_g, L.items_w, L.off_g, L.off_w,
                       n_rows, L.item_row);
    LSTC_LAUNCHED("k_sup_item_rows");
    LSTC_TRY(fill_u32(L.ticket, 0, 2, st));
    const uint32_t grid = (uint32_t)device_cus() * 4;  // persistent: four workgroups per CU (LDS)
    if (m < (1ull << 30))  // column positions as 32-bit byte offsets
        hipLaunchKernelGGL((k_support_split<true, kSupSplitWaves>), dim3(grid), dim3(kSupSplitWaves * 64), 0, st, d_off,
                           d_col, d_in_off,
                           d_in_col, n_rows, n_owned, (const uint64_t *)L.off_g, (const uint64_t *)L.off_w,
                           (const uint32_t *)L.item_row, L.ticket, d_support, (unsigned long long *)d_per_vertex);
    else
        hipLaunchKernelGGL((k_support_split<false, kSupSplitWaves>), dim3(grid), dim3(kSupSplitWaves * 64), 0, st,
                           d_off, d_col, d_in_off,
                           d_in_col, n_rows, n_owned, (const uint64_t *)L.off_g, (const uint64_t *)L.off_w,
                           (const uint32_t *)L.item_row, L.ticket, d_support, (unsigned long long *)d_per_vertex);
    LSTC_LAUNCHED("k_support_split");
    return LSTC_OK;
}
