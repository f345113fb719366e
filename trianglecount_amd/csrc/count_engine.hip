// count_engine.hip -- persistent triangle-counting engine (gfx950, wave64).
//
// Count of a graph = sum over owned rows u of sum_{w in N+(u)} |N+(u) n N+(w)|
// (src/WithoutPartition/ShareKernelTCV62_update_v1.cu:43-191 summed as at
// :271-277; owned rows only as in src/WithPartition/without_sorting_tc.cu:400).
//
// Every arc (u, w) contributes |N+(u) n N+(w)|, found by hashing one list into
// LDS and streaming the other through it.  A work item hashes one row r:
//   kind A (source side): N+(r) hashed; the lists N+(w), w in N+(r), streamed.
//   kind B (target side): N+(r) hashed; the lists N+(u) of in-neighbours u
//            (from the graph's in-lists) streamed.
// With in-lists present ("split" counting) kind A keeps only arcs with
// d+(w) < d+(r) and kind B covers the rest (the in-lists hold exactly the arcs
// u -> w with d+(u) <= d+(w)), so each arc streams its shorter list.  Without
// in-lists every arc is a kind A arc.
//
// Plan (k_plan_rows): every planned row with d+ >= 2 emits its kind A item
// (source side; owned rows only) and kind B items (one kInChunk-source chunk of
// its in-list each), binned by the hashed row's d+ into one queue, largest bin
// first; rows with d+ <= tiny_max go to the tiny region
// instead.  Items of rows above small_max form "big" tickets (kBigGroup items each);
// smaller ones form "small" tickets of small_group items (by default chosen by
// the planner from the small items per workgroup).  k_count reads it from ctl.
//
// Count (k_count, one persistent launch): each workgroup of WAVES waves (2, 4 or
// 8: the workgroup size is a tuning knob) takes
// tickets from an atomic head.
//   big item  : the whole workgroup builds one LDS hash table (up to 1024 slots
//               per wave: 4096 with 4 waves; load <= 1/2, longer rows are tiled)
//               and the waves split the 64-segment windows round-robin;
//   small item: one wave, a private 1024-slot slice of the same LDS.
// Walk of one window: one segment per lane (id, start, length), inclusive scan
// of lengths, empty segments compacted out; then the flat sequence of entries
// is consumed in rounds of UNROLL chunks of 64.  The segment ends inside a
// round are marked in LDS once, then every chunk finds each lane's segment with
// one LDS read + ballot + popcount (a chunk inside one segment takes its base
// from a scalar).  Rounds are software-pipelined: round r+1 is addressed and
// loaded (global loads, never FLAT) before round r is probed, lockstep over its
// UNROLL chunks.  Hits are counted per lane and reduced across the wave per
// item; the workgroup issues one uint64 atomic per graph it counted.
#include <mutex>
#include <unordered_map>

#include "count_engine.h"
#include "segwalk.h"
#include "wave.h"

namespace lstc {

static constexpr uint32_t kEmpty = 0xFFFFFFFFu;
static constexpr uint32_t kNoId = 0xFFFFFFFEu;  // never a vertex id (ids < 2^32 - 2), never kEmpty
static constexpr uint32_t kSliceLog2 = 10;
static constexpr uint32_t kSlice = 1u << kSliceLog2;  // wave-private slice (small items)
// LDS table of a workgroup: one slice per wave; a big item uses all of it as one
// table (2^big_log2 slots), filled with at most half as many keys per tile.
template <int WAVES>
constexpr uint32_t big_log2() { return kSliceLog2 + (WAVES >= 8 ? 3 : WAVES >= 4 ? 2 : 1); }
template <int WAVES>
constexpr uint32_t tab_slots() { return 1u << big_log2<WAVES>(); }
static constexpr int kBins = 6;                       // hash-path bins
static constexpr int kTinyBin = kBins;                // the lane-per-arc rows
static constexpr int kSlots = kBins + 1;              // queue regions
static constexpr uint32_t kBigGroup = 2;              // big items per ticket

// ctl (uint32) layout
enum { C_CNT = 0, C_HEAD = 12, C_NBIG = 13, C_NSMALL = 14, C_NTICK = 15, C_NTINY = 16, C_TINY0 = 17, C_NSMALLT = 18,
       C_SGROUP = 19, C_WORDS = 20 };

// work item: row (bits 0-31) | chunk (32-51) | kind (52) | graph (53-63).
// A row's in-degree is below 2^32 (m < 2^32), so its chunk index fits 20 bits.
__device__ __forceinline__ unsigned long long make_item(uint32_t g, uint32_t kind, uint32_t chunk, uint32_t r) {
    return ((unsigned long long)g << 53) | ((unsigned long long)kind << 52) | ((unsigned long long)chunk << 32) | r;
}
struct Item {
    uint32_t row, chunk, kind, g;
};
__device__ __forceinline__ Item decode(unsigned long long it) {
    return Item{(uint32_t)it, (uint32_t)(it >> 32) & 0xFFFFFu, (uint32_t)(it >> 52) & 1u, (uint32_t)(it >> 53)};
}

TuneR resolve_tune(const lstc_tune *t) {
    // defaults: the fastest point of a (table_scale, small_max, grid_per_cu, unroll)
    // sweep at RMAT scales 16, 21 and 23 on MI355X (table_scale 6: rows of up to
    // 170 keys keep a wave-private table)
    TuneR r{6, 256, 0, 0, 4, kTinyMax, 4};  // small_group 0: chosen by the planner
    if (t) {
        if (t->tiny_max) r.tiny_max = t->tiny_max == 1 ? 0u : t->tiny_max;  // 1 = off
        if (t->table_scale) r.table_scale = t->table_scale;
        if (t->small_max) r.small_max = t->small_max;
        if (t->small_group) r.small_group = t->small_group;
        if (t->grid_per_cu) r.grid_per_cu = t->grid_per_cu;
        if (t->unroll) r.unroll = t->unroll;
        if (t->waves) r.waves = t->waves;
    }
    r.table_scale = hmax(2u, hmin(16u, r.table_scale));
    // a small row must fit its slice at the chosen load factor
    r.small_max = hmax(64u, hmin(kSlice / r.table_scale, r.small_max));
    if (r.small_group) r.small_group = hmax(4u, hmin(256u, r.small_group));
    r.unroll = r.unroll >= 8 ? 8 : 4;
    r.waves = r.waves >= 8 ? 8 : r.waves >= 4 ? 4 : 2;
    // default: 32 waves per CU (k_count fits them in LDS and VGPRs), 8 workgroups of 4 waves
    if (r.grid_per_cu == 0) r.grid_per_cu = 32 / r.waves;
    r.grid_per_cu = hmax(1u, hmin(32u, r.grid_per_cu));
    r.tiny_max = hmin(kTinyMax, r.tiny_max);
    if (r.tiny_max < 2) r.tiny_max = 0;
    return r;
}

// A launch of few arcs has few items per wave: rows above 128 keys then
// share a workgroup table, so their lists spread over four waves (RMAT s16,
// 0.9M arcs: 20% faster; no change from 2^23 arcs up, where the default keeps
// rows of up to 1024 / table_scale keys wave-private).  Only a default
// small_max is capped; an explicit one is kept.
TuneR effective_tune(const lstc_tune *t, uint64_t total_m) {
    TuneR r = resolve_tune(t);
    if ((!t || !t->small_max) && total_m < (1ull << 23) && r.small_max > 128) r.small_max = 128;
    return r;
}

lstc_status validate_graph(const HostGraph &h, const char *who) {
    if (h.n_owned > h.n_rows) return fail(LSTC_ERR_INVALID, "%s: n_owned %u > n_rows %u", who, h.n_owned, h.n_rows);
    if (h.n_rows == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "%s: n_rows must be < 2^32-1", who);
    if (h.m >= 0xFFFFF000ull)
        return fail(LSTC_ERR_INVALID, "%s: m = %llu arcs exceeds 2^32-4096", who, (unsigned long long)h.m);
    if (h.n_owned && (!h.off || (h.m && !h.col))) return fail(LSTC_ERR_INVALID, "%s: NULL CSR arrays", who);
    if ((h.in_off == nullptr) != (h.in_col == nullptr))
        return fail(LSTC_ERR_INVALID, "%s: give both in-list arrays or neither", who);
    if ((h.col_seg == nullptr) != (h.in_seg == nullptr))
        return fail(LSTC_ERR_INVALID, "%s: give both segment-word arrays or neither", who);
    if (h.col_seg && !h.in_off) return fail(LSTC_ERR_INVALID, "%s: segment words without in-lists", who);
    if (h.scol && (!h.col_seg || h.scol_len >= 0xFFFFF000ull))
        return fail(LSTC_ERR_INVALID, "%s: a streaming copy needs segment words and < 2^32-4096 entries", who);
    return LSTC_OK;
}

struct Bounds {
    uint32_t hi[kBins];  // bin b holds d+ in (hi[b-1], hi[b]], bin 0 starts at 2
    uint32_t tiny_max;   // rows of graphs with in-lists and 2 <= d+ <= tiny_max go to the tiny region
};

// UNIFORM: g is the same in every lane, so the entry is read through the
// scalar cache into scalar registers.
template <bool UNIFORM = false>
__device__ __forceinline__ GraphDev graph_at(const GraphSet &gs, uint32_t g) {
    if (!gs.table) return gs.one;
    if (UNIFORM) {
        const LSTC_CONST_AS GraphDev *q = (const LSTC_CONST_AS GraphDev *)(gs.table + g);
        GraphDev d;
        d.off = q->off;
        d.col = q->col;
        d.per_row = q->per_row;
        d.in_off = q->in_off;
        d.in_col = q->in_col;
        d.n_rows = q->n_rows;
        d.n_owned = q->n_owned;
        d.col_seg = q->col_seg;
        d.in_seg = q->in_seg;
        d.scol = q->scol;
        return d;
    }
    const LSTC_GLOBAL_AS GraphDev *q = (const LSTC_GLOBAL_AS GraphDev *)(gs.table + g);
    GraphDev d;
    d.off = q->off;
    d.col = q->col;
    d.per_row = q->per_row;
    d.in_off = q->in_off;
    d.in_col = q->in_col;
    d.n_rows = q->n_rows;
    d.n_owned = q->n_owned;
    d.scol = q->scol;
    return d;
}

__device__ __forceinline__ uint32_t table_log2_for(uint32_t keys, uint32_t scale, uint32_t max_log2) {
    uint32_t l = 6;
    while (l < max_log2 && (1u << l) < scale * keys) ++l;
    return l;
}

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// ------------------------------------------------------------------ planning
__global__ void k_plan_init(uint32_t *ctl, unsigned long long *counts, uint32_t G) {
    for (uint32_t i = threadIdx.x; i < C_WORDS; i += blockDim.x) ctl[i] = 0;
    for (uint32_t i = threadIdx.x; i < G; i += blockDim.x) counts[i] = 0;
}

// The plan splits the owned rows of the launch (graph after graph) into
// kPlanWaves contiguous ranges, one per wave of the plan grid.  Pass 0 counts
// each range's items per bin into wcnt[slot][wave]; k_plan_scan turns that into
// exclusive queue positions (slot-major: bins 5,4,3 = big, largest first, then
// 2,1,0 = small); pass 1 writes the items, in row order within a bin and range.
// No atomics: the queue is the same on every run.
static constexpr uint32_t kPlanBlocks = 1024;
static constexpr uint32_t kPlanWaves = kPlanBlocks * 4;
static constexpr uint32_t kPlanEntries = kSlots * kPlanWaves;

// queue region of a bin: hash bins 5..0 in slots 0..5, the tiny rows in slot 6
__device__ __forceinline__ int slot_of(int b) { return b < kBins ? kBins - 1 - b : kTinyBin; }

template <int PASS>
__global__ __launch_bounds__(256) void k_plan_rows(GraphSet gs, Bounds bd, uint32_t *__restrict__ wcnt,
                                                   unsigned long long *__restrict__ queue) {
    const uint32_t lane = lane_id();
    const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t total = gs.total_rows;
    const uint64_t chunk = (total + kPlanWaves - 1) / kPlanWaves;
    const uint64_t r_begin = umin64(total, (uint64_t)wave * chunk), r_end = umin64(total, r_begin + chunk);
    uint32_t run[kSlots];
#pragma unroll
    for (int i = 0; i < kSlots; ++i) run[i] = PASS == 1 ? wcnt[slot_of(i) * kPlanWaves + wave] : 0u;
    for (uint64_t r0 = r_begin; r0 < r_end; r0 += 64) {
        const uint64_t r = r0 + lane;
        int b = -1;
        uint32_t g = 0, u = 0, n_a = 0, n_b = 0, tiny_kind = 0;
        if (r < r_end) {
            if (gs.table) {  // last g with row_base[g] <= r
                uint32_t lo = 0, hi = gs.G;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (gload(gs.row_base + mid) <= r) lo = mid; else hi = mid;
                }
                g = lo;
                u = (uint32_t)(r - gload(gs.row_base + g));
            } else {
                u = (uint32_t)r;
            }
            const GraphDev G = graph_at(gs, g);
            const bool owned = u < G.n_owned;  // rows past n_owned are visited only for split counting
            const uint64_t k = gload(G.off + u + 1) - gload(G.off + u);
            const uint64_t kb = G.in_off ? gload(G.in_off + u + 1) - gload(G.in_off + u) : 0;
            if (k >= 2 && G.in_off && k <= bd.tiny_max) {
                // one entry; kind 1 = in side only (a halo row owns no out side)
                if (owned || kb) {
                    b = kTinyBin;
                    tiny_kind = owned ? 0u : 1u;
                }
            } else if (k >= 2) {
                const uint32_t sb = in_chunk(kb);
                n_a = owned ? 1u : 0u;
                n_b = (uint32_t)((kb + sb - 1) / sb);
                if (n_a + n_b) {
                    b = kBins - 1;
                    for (int i = kBins - 2; i >= 0; --i)
                        if (k <= bd.hi[i]) b = i;
                }
            }
        }
        const uint32_t n_items = b == kTinyBin ? 1u : n_a + n_b;
#pragma unroll
        for (int i = 0; i < kSlots; ++i) {
            if (ballot(b == i) == 0) continue;
            const uint32_t c = (b == i) ? n_items : 0u;
            if (PASS == 0) {
                run[i] += c;
            } else {
                const uint32_t incl = wave_incl_scan(c, lane);
                if (b == i) {
                    unsigned long long *q = queue + run[i] + (incl - c);
                    if (b == kTinyBin) {
                        q[0] = make_item(g, tiny_kind, 0, u);
                    } else {
                        if (n_a) *q++ = make_item(g, 0, 0, u);
                        for (uint32_t j = 0; j < n_b; ++j) q[j] = make_item(g, 1, j, u);
                    }
                }
                run[i] += read_lane(incl, 63);
            }
        }
    }
    if (PASS == 0) {
#pragma unroll
        for (int i = 0; i < kSlots; ++i) {
            const uint32_t sum = wave_sum32(run[i]);
            if (lane == 0) wcnt[slot_of(i) * kPlanWaves + wave] = sum;
        }
    }
}

// One workgroup of 1024: exclusive scan of wcnt in place (kPlanEntries values,
// 24 per thread), then the ticket counts.  Item totals stay below 2^32
// (checked on the host against the queue capacity).
// small_group 0: the planner picks the small items per ticket, the largest
// power of two up to (small items per workgroup) / 32, within [16, 64]: about
// 32-64 small tickets per workgroup (measured on RMAT s21 and s23: fewer,
// larger tickets leave a ragged tail, more of them queue on the ticket head).
__global__ __launch_bounds__(1024) void k_plan_scan(uint32_t *__restrict__ wcnt, uint32_t *__restrict__ ctl,
                                                    uint32_t small_group, uint32_t grid, uint32_t waves) {
    constexpr uint32_t kPer = kPlanEntries / 1024;
    static_assert(kPlanEntries % 1024 == 0, "plan entries per thread");
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_slot[kSlots + 1];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t v[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        v[i] = wcnt[tid * kPer + i];
        sum += v[i];
    }
    const uint32_t incl = wave_incl_scan(sum, lane);
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wv; ++w) before += s_wave[w];
    uint32_t run = before + incl - sum;
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        const uint32_t e = tid * kPer + i;
        if (e % kPlanWaves == 0) s_slot[e / kPlanWaves] = run;
        wcnt[e] = run;
        run += v[i];
    }
    if (tid == 1023) s_slot[kSlots] = run;
    __syncthreads();
    if (tid == 0) {
        // slots 0..2 hold bins 5,4,3 (big), slots 3..5 bins 2,1,0 (small), slot 6 the tiny rows
        const uint32_t n_big = s_slot[3] - s_slot[0];
        const uint32_t n_small = s_slot[6] - s_slot[3];
        const uint32_t n_tiny = s_slot[7] - s_slot[6];
        for (int b = 0; b < kBins; ++b) ctl[C_CNT + b] = s_slot[kBins - b] - s_slot[kBins - 1 - b];
        uint32_t sg = small_group;
        if (sg == 0) {
            const uint32_t per_wg = n_small / (grid ? grid : 1u) / 32u;
            sg = 16;
            while (sg < 64 && 2 * sg <= per_wg) sg *= 2;
        }
        ctl[C_SGROUP] = sg;
        const uint32_t small_t = (n_small + sg - 1) / sg;
        const uint32_t tiny_t = (n_tiny + waves * 64 - 1) / (waves * 64);  // 64 rows per wave
        ctl[C_NBIG] = n_big;
        ctl[C_NSMALL] = n_small;
        ctl[C_NSMALLT] = small_t;
        ctl[C_NTINY] = n_tiny;
        ctl[C_TINY0] = s_slot[6];
        ctl[C_NTICK] = (n_big + kBigGroup - 1) / kBigGroup + small_t + tiny_t;
        ctl[C_HEAD] = 0;
    }
}

// ------------------------------------------------------------------ table build
// A key whose home bucket is full moves on; its home is then marked in the
// overflow bitmap `ovf` (one bit per bucket).  table_mark afterwards records
// the mark in the bucket itself by the order of its first two slots (an
// overflowed bucket is full: slot 0 > slot 1; any other: slot 0 < slot 1, or
// an empty slot 1), so a probe learns from its one bucket read whether a key
// of that home can lie further on.
__device__ __forceinline__ void table_insert(uint32_t *tab, uint32_t *ovf, uint32_t tlog2,
                                             const uint32_t *__restrict__ col, uint64_t b0, uint32_t nk, uint32_t tid,
                                             uint32_t nthreads) {
    const uint32_t qmask = (1u << (tlog2 - 2)) - 1, qshift = 34 - tlog2;
    for (uint32_t j = tid; j < nk; j += nthreads) {
        const uint32_t key = gload(col + b0 + j);
        const uint32_t home = hquad(key, qshift);
        uint32_t h = home;
        // A slot is tried only once the ones before it in the bucket are taken.
        // At most 1/2 of the slots are filled, so a free slot lies within one
        // lap of the buckets; the lap bound makes the loop finite by
        // construction as well.
        for (uint32_t step = 0; step <= qmask; ++step) {
            uint32_t *b = tab + 4 * h;
            if (atomicCAS(b, kEmpty, key) == kEmpty || atomicCAS(b + 1, kEmpty, key) == kEmpty ||
                atomicCAS(b + 2, kEmpty, key) == kEmpty || atomicCAS(b + 3, kEmpty, key) == kEmpty)
                break;
            h = (h + 1) & qmask;
        }
        if (h != home) atomicOr(ovf + (home >> 5), 1u << (home & 31u));
    }
}

// Orders the first two slots of every bucket by its overflow mark (the keys
// stay in the bucket, so every chain and membership is unchanged).
__device__ __forceinline__ void table_mark(uint32_t *tab, const uint32_t *ovf, uint32_t tlog2, uint32_t tid,
                                           uint32_t nthreads) {
    for (uint32_t b = tid; b < (1u << (tlog2 - 2)); b += nthreads) {
        const uint2 v = *(const uint2 *)(tab + 4 * b);
        const bool over = (ovf[b >> 5] >> (b & 31u)) & 1u;
        if (over ? v.x < v.y : v.x > v.y) *(uint2 *)(tab + 4 * b) = uint2{v.y, v.x};
    }
}

// ------------------------------------------------------------------ walk
// Probes one round through tab: linear probing over 4-slot buckets, the
// first bucket of all UNROLL chunks in lockstep, one 16-byte LDS read per
// chunk.  A lane is done once it hits or reads a bucket whose last slot is
// empty (no chain continues past such a bucket; tables are at most half full,
// so every chain has one).  Positions past the window's end carry an id that
// no table holds (kNoId: ids are below 2^32 - 2).
// Each compare is balloted on its own (a ballot of one compare is the
// compare's lane mask; a ballot of a combined condition would go through a
// VGPR), the masks are combined in scalar registers, and the hit mask is added
// to the per-lane counter as the carry-in of one vector add.  The few lanes
// whose chain goes on (about one lane in a hundred at the default table scale)
// are queued (their ids) in the window's segment table, which the rounds no
// longer read (bases and ends are in registers), and probed 64 at a time when
// the queue fills and at the window's end.
__device__ __forceinline__ uint64_t bucket_hits(const uint4 &v, uint32_t x) {
    return ballot(v.x == x) | ballot(v.y == x) | ballot(v.z == x) | ballot(v.w == x);
}

// Probes the queued lanes of a window (ids qx[i], i < qn, whose home bucket
// was full without a hit) as one chunk, from the bucket after their home to
// the end of their chains.
__device__ __forceinline__ void drain_queue(const uint4 *buckets, uint32_t qmask, uint32_t qshift, uint32_t lane,
                                            uint32_t &lane_hits, const uint32_t *qx, uint32_t qn) {
    const bool act = lane < qn;
    const uint32_t xv = act ? qx[lane] : kNoId;
    uint32_t q = (hquad(xv, qshift) + 1u) & qmask;
    uint64_t live = ballot(act);
    // Every chain ends at a hit or at a bucket whose last slot is empty: slots
    // fill in order, so a bucket with its last slot taken is full, and with at
    // most 1/2 of the slots filled at least half of the buckets end chains.
    // A chain therefore ends within one lap; the lap bound keeps the loop
    // finite by construction as well.
    for (uint32_t step = 0; live != 0 && step < qmask; ++step) {
        const uint4 w = buckets[q];
        const uint64_t H = bucket_hits(w, xv) & live;
        add_lane_bit(lane_hits, H);
        live &= ~(H | ballot(w.w == kEmpty));
        q = (q + 1) & qmask;
    }
}

template <int UNROLL, bool FULL>
__device__ __forceinline__ void probe_round(const uint32_t (&xin)[UNROLL], uint32_t B, uint32_t total, const uint32_t *tab,
                                            uint32_t qmask, uint32_t qshift, uint32_t lane, uint32_t &lane_hits,
                                            uint32_t *qx, uint32_t &qn) {
    const uint4 *buckets = (const uint4 *)tab;
    const uint64_t active = ballot(true);
    uint32_t x[UNROLL], p[UNROLL];
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) {
        x[k] = FULL || B + 64u * k + lane < total ? xin[k] : kNoId;
        p[k] = hquad(x[k], qshift);
    }
    uint4 v[UNROLL];
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) v[k] = buckets[p[k]];
    // lanes whose chain goes on (bits of inactive lanes may be set here; they
    // are masked off below)
    uint64_t going[UNROLL], any = 0;
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) {
        const uint64_t H = bucket_hits(v[k], x[k]);
        add_lane_bit(lane_hits, H);
        going[k] = ~(H | ballot(v[k].x <= v[k].y));  // no hit and the home bucket overflowed
        any |= going[k];
    }
    if ((any & active) == 0) return;
    // queue them (id, next bucket); a full queue is probed first
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) {
        const uint64_t g = going[k] & active;
        if (g != 0) {
            const uint32_t c = (uint32_t)__popcll(g);
            if (qn + c > 64u) {
                drain_queue(buckets, qmask, qshift, lane, lane_hits, qx, qn);
                qn = 0;
            }
            if (lane_in(g)) qx[qn + mask_rank(g)] = x[k];
            qn += c;
        }
    }
}

// Streams the lists N+(s) of the segment ids s = seg[j], j in [w0, nseg) step
// wstep (windows of 64), through tab.  A segment whose list is not shorter than
// `below` is skipped (below = 0: none skipped); ids >= n_rows have no list.
// With segment words (segw[j] = start | length << 32 of N+(seg[j])) a window
// reads one coalesced word per segment instead of two row offsets at a random
// place, and the next window's words are loaded while this one streams.
// Rounds are software-pipelined over two buffers: the loads of round r+1 are
// issued before round r is probed, so a wave keeps up to 2*UNROLL chunk loads
// in flight.
template <int UNROLL, bool NARROW>
__device__ __forceinline__ uint32_t walk(const uint64_t *__restrict__ off, const uint32_t *__restrict__ col,
                                                   uint32_t n_rows, const uint32_t *__restrict__ seg,
                                                   const uint64_t *__restrict__ segw, uint32_t nseg_all,
                                                   uint32_t below, uint32_t w0, uint32_t wstep, const uint32_t *tab,
                                                   uint32_t tlog2, uint32_t *s_base, uint32_t *s_end,
                                                   uint32_t lane) {
    constexpr uint32_t RB = 64u * UNROLL;
    const uint32_t qmask = (1u << (tlog2 - 2)) - 1, qshift = 34 - tlog2;
    uint32_t lane_hits = 0;
    // the next window's segment words are loaded while this window streams
    uint64_t sw_next = segw && w0 + lane < nseg_all ? gload(segw + w0 + lane) : 0;
    for (uint32_t j0 = w0; j0 < nseg_all; j0 += wstep) {
        const uint32_t j = j0 + lane;
        uint32_t s = 0, len = 0;
        if (segw) {
            const uint64_t sw = sw_next;  // one coalesced word: start | length << 32 (0 past the end)
            s = (uint32_t)sw;
            len = (uint32_t)(sw >> 32);
            if (below && len >= below) len = 0;  // that arc is counted from the other side
            sw_next = j + wstep < nseg_all ? gload(segw + j + wstep) : 0;
        } else if (j < nseg_all) {
            const uint32_t w = gload(seg + j);
            if (w < n_rows) {
                const uint64_t s0 = gload(off + w);
                s = (uint32_t)s0;
                len = (uint32_t)(gload(off + w + 1) - s0);
            }
            if (below && len >= below) len = 0;
        }
        const uint32_t incl = wave_incl_scan(len, lane);
        const uint32_t total = uniform(gather(incl, 63));
        if (total == 0) continue;
        const uint64_t nzm = ballot(len != 0);
        const uint32_t nseg = (uint32_t)__popcll(nzm);
        if (len != 0) {
            const uint32_t slot = mask_rank(nzm);
            s_base[slot] = s - (incl - len);
            s_end[slot] = incl;
        }
        const uint32_t mybase = s_base[lane];
        const uint32_t myend = lane < nseg ? s_end[lane] : kEmpty;
        uint32_t xa[UNROLL], xb[UNROLL];
        uint32_t qn = 0;  // queued lanes (in s_base from here on)
        // Each round's indices come from one of three cases (inside one
        // segment, crossed by segment ends, the window's partial last round);
        // the loads and the probe follow in one common path, so the two
        // buffers keep their registers and the wait for the probed buffer
        // never covers the loads just issued.
        uint32_t idx[UNROLL];
        {
            const uint32_t sg = first_segment(0, myend);
            if (RB > total)
                index_walk<UNROLL, false, NARROW>(0, total, mybase, myend, sg, lane, idx);
            else if (one_segment<UNROLL>(0, myend))
                index_single<UNROLL, NARROW>((uint32_t)__builtin_amdgcn_readlane((int)mybase, (int)sg), lane, idx);
            else
                index_walk<UNROLL, true, NARROW>(0, total, mybase, myend, sg, lane, idx);
        }
        load_round<UNROLL, NARROW>(col, idx, xa);
        uint32_t B = 0;
        for (; B + 3 * RB <= total; B += 2 * RB) {
            // rounds B + RB and B + 2 * RB are full
            uint32_t sg = first_segment(B + RB, myend);
            if (one_segment<UNROLL>(B + RB, myend))
                index_single<UNROLL, NARROW>((uint32_t)__builtin_amdgcn_readlane((int)mybase, (int)sg) + B + RB, lane, idx);
            else
                index_walk<UNROLL, true, NARROW>(B + RB, total, mybase, myend, sg, lane, idx);
            load_round<UNROLL, NARROW>(col, idx, xb);
            probe_round<UNROLL, true>(xa, B, total, tab, qmask, qshift, lane, lane_hits, s_base, qn);
            sg = first_segment(B + 2 * RB, myend);
            if (one_segment<UNROLL>(B + 2 * RB, myend))
                index_single<UNROLL, NARROW>((uint32_t)__builtin_amdgcn_readlane((int)mybase, (int)sg) + B + 2 * RB, lane,
                                     idx);
            else
                index_walk<UNROLL, true, NARROW>(B + 2 * RB, total, mybase, myend, sg, lane, idx);
            load_round<UNROLL, NARROW>(col, idx, xa);
            probe_round<UNROLL, true>(xb, B + RB, total, tab, qmask, qshift, lane, lane_hits, s_base, qn);
        }
        // Tail: round B is in xa; rounds B + RB and B + 2 * RB may follow
        // (B + 3 * RB > total, so the last one is partial).
        if (B + RB >= total) {
            probe_round<UNROLL, false>(xa, B, total, tab, qmask, qshift, lane, lane_hits, s_base, qn);
        } else {
            index_walk<UNROLL, false, NARROW>(B + RB, total, mybase, myend, first_segment(B + RB, myend), lane, idx);
            load_round<UNROLL, NARROW>(col, idx, xb);
            probe_round<UNROLL, true>(xa, B, total, tab, qmask, qshift, lane, lane_hits, s_base, qn);
            if (B + 2 * RB < total) {
                index_walk<UNROLL, false, NARROW>(B + 2 * RB, total, mybase, myend, first_segment(B + 2 * RB, myend), lane,
                                          idx);
                load_round<UNROLL, NARROW>(col, idx, xa);
                probe_round<UNROLL, true>(xb, B + RB, total, tab, qmask, qshift, lane, lane_hits, s_base, qn);
                probe_round<UNROLL, false>(xa, B + 2 * RB, total, tab, qmask, qshift, lane, lane_hits, s_base, qn);
            } else {
                probe_round<UNROLL, false>(xb, B + RB, total, tab, qmask, qshift, lane, lane_hits, s_base, qn);
            }
        }
        // the next window rewrites the segment tables
        if (qn != 0) drain_queue((const uint4 *)tab, qmask, qshift, lane, lane_hits, s_base, qn);
    }
    // an item streams at most m < 2^32 entries, so its hits fit 32 bits
    return wave_sum32(lane_hits);
}

// The segment source of a work item: kind A walks N+(r) with the split filter,
// kind B walks one chunk of r's in-list unfiltered.
struct Segs {
    const uint32_t *ids;
    const uint64_t *seg;  // the ids' segment words, or null
    uint32_t n;
    uint32_t below;
};

__device__ __forceinline__ Segs item_segs(const GraphDev &G, uint32_t kind, uint32_t chunk, uint32_t r, uint64_t rb,
                                          uint32_t K) {
    if (kind == 0) return Segs{G.col + rb, G.col_seg ? G.col_seg + rb : nullptr, K, G.in_off ? K : 0u};
    const uint64_t ib = sload(G.in_off + r), ie = sload(G.in_off + r + 1);
    const uint32_t sb = in_chunk(ie - ib);
    const uint64_t c0 = ib + (uint64_t)chunk * sb;
    const uint64_t c1 = (c0 + sb < ie) ? c0 + sb : ie;
    return Segs{G.in_col + c0, G.in_seg ? G.in_seg + c0 : nullptr, (uint32_t)(c1 - c0), 0u};
}

// ------------------------------------------------------------------ tiny rows
// Rows with 2 <= d+ <= tiny_max of a graph with in-lists.  Under the split rule
// every arc such a row r owns joins two lists of at most d+(r) entries:
//   out side: arcs (r, x) with d+(x) < d+(r), from N+(r);
//   in side : arcs (u, r) from r's in-list (d+(u) <= d+(r)).
// So each arc is one lane's job: compare the two lists, at most kTinyMax^2
// equality tests, no hash table and no probing.  The wave's lanes hold up to 64
// rows of graph G (mine); their lists are staged once in the wave's LDS slice
// (row of lane i at tab[i*kTinyMax ...]), then their out-sides and in-sides are
// walked flat, 64 arcs per step, with the same end-marking as the hash walk.
// Each lane loads the other list of its arc into registers.
__device__ __forceinline__ uint32_t tiny_walk(const GraphDev &G, bool mine, bool out_side, uint32_t r, uint32_t *tab,
                                              uint32_t *s_base, uint32_t *s_end, uint8_t *s_src, uint8_t *flag,
                                              uint32_t lane) {
    uint32_t ro = 0, rd = 0;  // this lane's row: start and length of N+(r)
    if (mine) {
        const uint64_t o0 = gload(G.off + r);
        ro = (uint32_t)o0;
        rd = (uint32_t)(gload(G.off + r + 1) - o0);
#pragma unroll
        for (uint32_t j = 0; j < kTinyMax; ++j)
            if (j < rd) tab[lane * kTinyMax + j] = gload(G.col + ro + j);
    }
    uint32_t lane_hits = 0;
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
        uint32_t s = 0, len = 0;
        if (mine) {
            if (side == 0) {
                s = ro;
                len = out_side ? rd : 0u;
            } else {
                const uint64_t i0 = gload(G.in_off + r);
                s = (uint32_t)i0;
                len = (uint32_t)(gload(G.in_off + r + 1) - i0);
            }
        }
        const uint32_t *list = side == 0 ? G.col : G.in_col;
        const uint32_t incl = wave_incl_scan(len, lane);
        const uint32_t total = uniform(gather(incl, 63));
        if (total == 0) continue;
        const uint64_t nzm = ballot(len != 0);
        const uint32_t nseg = (uint32_t)__popcll(nzm);
        if (len != 0) {
            const uint32_t slot = mask_rank(nzm);
            s_base[slot] = s - (incl - len);
            s_end[slot] = incl;
            s_src[slot] = (uint8_t)lane;
        }
        const uint32_t mybase = s_base[lane];
        const uint32_t myend = lane < nseg ? s_end[lane] : kEmpty;
        const uint32_t mysrc = s_src[lane];
        for (uint32_t B = 0; B < total; B += 64) {
            uint32_t sg = (uint32_t)__popcll(ballot(myend <= B));
            const bool mk = myend > B && myend < B + 64;
            if (mk) flag[myend - B] = 1;
            const uint64_t M = ballot(flag[lane] != 0);
            if (mk) flag[myend - B] = 0;
            if (M != 0) sg += (uint32_t)__popcll(M & ((2ull << lane) - 1ull));
            sg &= 63u;  // lanes past the end take some owner; their arc is not counted
            const uint32_t base = gather(mybase, sg), src = gather(mysrc, sg);
            const uint32_t t = B + lane;
            const uint32_t x = gload(t < total ? list + (uint32_t)(base + t) : list);
            const uint32_t ard = gather(rd, src);
            // x's list: with segment words, the arc's own word (coalesced) gives its start in the
            // streamed array (the row-aligned copy, or col) and d+(x); else two row offsets
            const uint64_t *words = side == 0 ? G.col_seg : G.in_seg;
            const uint32_t *xl = words ? G.scol : G.col;
            uint32_t xo, xd;
            if (words) {
                const uint64_t w = gload(t < total ? words + (uint32_t)(base + t) : words);
                xo = (uint32_t)w;
                xd = t < total ? (uint32_t)(w >> 32) : 0u;  // 0 for ids that are no row
            } else {
                const uint32_t xi = t < total ? x : 0u;
                const uint64_t xo0 = gload(G.off + xi);
                xo = (uint32_t)xo0;
                xd = (uint32_t)(gload(G.off + xi + 1) - xo0);
            }
            // the out side keeps d+(x) < d+(r); the in side holds only d+(x) <= d+(r)
            const bool live = t < total && xd != 0 && (side == 1 || xd < ard);
            uint32_t b[kTinyMax];
#pragma unroll
            for (uint32_t j = 0; j < kTinyMax; ++j)  // unconditional, in bounds (xl[0] when unused)
                b[j] = gload(live && j < xd ? xl + xo + j : xl);
            if (live) {
                const uint32_t *a = tab + src * kTinyMax;
                uint32_t c = 0;
#pragma unroll
                for (uint32_t i = 0; i < kTinyMax; ++i) {
                    if (i < ard) {
                        const uint32_t ai = a[i];
#pragma unroll
                        for (uint32_t j = 0; j < kTinyMax; ++j) c += (j < xd && ai == b[j]) ? 1u : 0u;
                    }
                }
                lane_hits += c;
            }
        }
    }
    return wave_sum32(lane_hits);
}

// ------------------------------------------------------------------ persistent count kernel
// NARROW: every graph of the launch has m <= 2^30 arcs (and a streaming copy of at
// most 2^30 entries).
// Occupancy target (the second launch bound is waves per SIMD): 8 with UNROLL 4,
// 6 with UNROLL 8, whatever the workgroup size.
template <int UNROLL, bool NARROW, int WAVES>
__global__ __launch_bounds__(WAVES * 64, UNROLL == 4 ? 8 : 6) void k_count(GraphSet gs, const unsigned long long *__restrict__ queue,
                                                       uint32_t *__restrict__ ctl,
                                                       unsigned long long *__restrict__ counts, uint32_t scale) {
    constexpr uint32_t kTab = tab_slots<WAVES>();
    __shared__ __align__(16) uint32_t s_tab[kTab];
    __shared__ uint32_t s_ovf[kTab / 128];  // overflow bits, one per 4-slot bucket
    __shared__ uint32_t s_base[WAVES][64];
    __shared__ uint32_t s_end[WAVES][64];
    __shared__ __align__(16) uint32_t s_flag[WAVES][16];  // 64 position flags per wave (tiny rows)
    __shared__ uint8_t s_src[WAVES][64];
    __shared__ uint32_t s_ticket;
    __shared__ unsigned long long s_acc[WAVES];
    __shared__ uint32_t s_g[WAVES];

    // item and graph fields are wave-uniform: they live in scalar registers
    const uint32_t tid = threadIdx.x, wv = uniform(tid >> 6), lane = tid & 63;
    if (lane < 16) s_flag[wv][lane] = 0;
    const uint32_t n_big = sload(ctl + C_NBIG), n_small = sload(ctl + C_NSMALL), n_tickets = sload(ctl + C_NTICK);
    const uint32_t n_bigt = (n_big + kBigGroup - 1) / kBigGroup;  // big tickets
    const uint32_t small_end = n_bigt + sload(ctl + C_NSMALLT);  // first tiny ticket
    const uint32_t small_group = sload(ctl + C_SGROUP);
    const uint32_t tiny0 = sload(ctl + C_TINY0), tiny_end = tiny0 + sload(ctl + C_NTINY);
    uint32_t cur_g = kEmpty;
    unsigned long long acc = 0;

    for (;;) {
        if (tid == 0) s_ticket = atomicAdd(ctl + C_HEAD, 1u);
        __syncthreads();
        const uint32_t t = uniform(s_ticket);
        __syncthreads();
        if (t >= n_tickets) break;
        if (t < n_bigt) {
          // big items (kBigGroup per ticket): the whole workgroup shares one table
          for (uint32_t bi = t * kBigGroup, be = min(bi + kBigGroup, n_big); bi < be; ++bi) {
            const Item itm = decode(sload(queue + bi));
            const uint32_t u = itm.row, chunk = itm.chunk, kind = itm.kind, g = itm.g;
            const GraphDev G = graph_at<true>(gs, g);
            const uint64_t rb = sload(G.off + u);
            const uint32_t K = (uint32_t)(sload(G.off + u + 1) - rb);
            const Segs sg = item_segs(G, kind, chunk, u, rb, K);
            unsigned long long hits = 0;
            constexpr uint32_t kBigTile = kTab / 2;  // keys per table fill of a big row
            for (uint32_t t0 = 0; t0 < K; t0 += kBigTile) {
                const uint32_t tk = min(K - t0, kBigTile);
                const uint32_t tl = table_log2_for(tk, scale, big_log2<WAVES>());
                for (uint32_t i = tid; i < (1u << tl); i += WAVES * 64) s_tab[i] = kEmpty;
                if (tid < kTab / 128) s_ovf[tid] = 0;
                __syncthreads();
                table_insert(s_tab, s_ovf, tl, G.col, rb + t0, tk, tid, WAVES * 64);
                __syncthreads();
                table_mark(s_tab, s_ovf, tl, tid, WAVES * 64);
                __syncthreads();
                hits += walk<UNROLL, NARROW>(G.off, sg.seg ? G.scol : G.col, G.n_rows, sg.ids, sg.seg, sg.n, sg.below, wv * 64, WAVES * 64, s_tab, tl,
                                     s_base[wv], s_end[wv], lane);
                __syncthreads();
            }
            if (g != cur_g) {
                if (cur_g != kEmpty && acc != 0 && lane == 0) atomicAdd(counts + cur_g, acc);
                cur_g = g;
                acc = 0;
            }
            acc += hits;
            if (G.per_row && lane == 0 && hits) atomicAdd(G.per_row + u, hits);
          }
        } else if (t >= small_end) {
            // tiny rows: 64 per wave, lane-per-arc; the lanes are split by graph
            const uint32_t e0 = tiny0 + (t - small_end) * (WAVES * 64) + wv * 64;
            const uint32_t e = e0 + lane;
            const bool valid = e < tiny_end;
            Item itm{0, 0, 0, 0};
            if (valid) itm = decode(gload(queue + e));
            uint64_t rem = ballot(valid);
            while (rem) {
                const uint32_t g = read_lane(itm.g, (uint32_t)__builtin_ctzll(rem));
                const bool mine = valid && itm.g == g;
                rem &= ~ballot(mine);
                const GraphDev G = graph_at<true>(gs, g);
                const uint32_t hits = tiny_walk(G, mine, mine && itm.kind == 0, itm.row, s_tab + wv * kSlice,
                                                s_base[wv], s_end[wv], s_src[wv], (uint8_t *)s_flag[wv], lane);
                if (g != cur_g) {
                    if (cur_g != kEmpty && acc != 0 && lane == 0) atomicAdd(counts + cur_g, acc);
                    cur_g = g;
                    acc = 0;
                }
                acc += hits;
            }
        } else {
            // small items: each wave takes every WAVES-th item of the group
            const uint32_t first = n_big + (t - n_bigt) * small_group;
            const uint32_t last = min(first + small_group, n_big + n_small);
            uint32_t *tab = s_tab + wv * kSlice;
            for (uint32_t q = first + wv; q < last; q += WAVES) {
                const Item itm = decode(sload(queue + q));
                const uint32_t u = itm.row, chunk = itm.chunk, kind = itm.kind, g = itm.g;
                const GraphDev G = graph_at<true>(gs, g);
                const uint64_t rb = sload(G.off + u);
                const uint32_t K = (uint32_t)(sload(G.off + u + 1) - rb);
                const Segs sg = item_segs(G, kind, chunk, u, rb, K);
                const uint32_t tl = table_log2_for(K, scale, kSliceLog2);
                uint32_t *ovf = s_ovf + wv * (kSlice / 128);
                for (uint32_t i = lane; i < (1u << tl); i += 64) tab[i] = kEmpty;
                if (lane < kSlice / 128) ovf[lane] = 0;
                table_insert(tab, ovf, tl, G.col, rb, K, lane, 64);
                table_mark(tab, ovf, tl, lane, 64);
                const unsigned long long hits =
                    walk<UNROLL, NARROW>(G.off, sg.seg ? G.scol : G.col, G.n_rows, sg.ids, sg.seg, sg.n, sg.below, 0, 64, tab, tl,
                                 s_base[wv], s_end[wv], lane);
                if (g != cur_g) {
                    if (cur_g != kEmpty && acc != 0 && lane == 0) atomicAdd(counts + cur_g, acc);
                    cur_g = g;
                    acc = 0;
                }
                acc += hits;
                if (G.per_row && lane == 0) G.per_row[u] = hits;
            }
        }
    }
    // one atomic per graph per workgroup for what the waves still hold
    if (lane == 0) { s_acc[wv] = acc; s_g[wv] = cur_g; }
    __syncthreads();
    if (tid == 0) {
        for (int i = 0; i < WAVES; ++i) {
            if (s_g[i] == kEmpty) continue;
            unsigned long long sum = s_acc[i];
            for (int j = i + 1; j < WAVES; ++j)
                if (s_g[j] == s_g[i]) { sum += s_acc[j]; s_g[j] = kEmpty; }
            if (sum) atomicAdd(counts + s_g[i], sum);
        }
    }
}

// ------------------------------------------------------------------ host side
// Workgroups of `kern` that fit on one CU at once (registers, LDS), cached per
// kernel; 0 when the runtime cannot tell.
static uint32_t resident_blocks(const void *kern, int threads) {
    static std::mutex mu;
    static std::unordered_map<const void *, uint32_t> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(kern);
    if (it != cache.end()) return it->second;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kern, threads, 0) != hipSuccess || blocks < 0) blocks = 0;
    (void)hipGetLastError();
    cache[kern] = (uint32_t)blocks;
    return (uint32_t)blocks;
}

// queue entries: per row one kind A item and at most 1 + in-degree / kInChunk kind B items
static uint64_t queue_cap(uint64_t total_rows, uint64_t total_m) { return 2 * total_rows + total_m / kInChunk + 1; }

size_t count_ws_bytes(uint32_t G, uint64_t total_rows, uint64_t total_m) {
    (void)G;
    Carver c(nullptr, 0);
    c.take<uint32_t>(64);
    c.take<uint32_t>(kPlanEntries);
    c.take<unsigned long long>(queue_cap(total_rows, total_m));
    return c.need();
}

lstc_status count_launch(const HostGraph *hosts, uint32_t G, const GraphDev *d_table, const uint64_t *d_row_base,
                         unsigned long long *d_counts, const lstc_tune *tune, void *ws, size_t ws_bytes,
                         hipStream_t st, bool timed) {
    if (G == 0) return LSTC_OK;
    if (G > kMaxGraphs) return fail(LSTC_ERR_INVALID, "count: at most %u graphs per launch", kMaxGraphs);
    if (!d_counts) return fail(LSTC_ERR_INVALID, "count: NULL counts");
    if (G > 1 && (!d_table || !d_row_base)) return fail(LSTC_ERR_INVALID, "count: batch needs device descriptors");
    uint64_t total = 0, total_m = 0;
    for (uint32_t g = 0; g < G; ++g) {
        total += plan_rows(hosts[g]);
        total_m += hosts[g].m;
    }
    const TuneR tr = effective_tune(tune, total_m);
    if (total >= 0xFFFFFFFFull) return fail(LSTC_ERR_INVALID, "count: more than 2^32-1 owned rows in one launch");
    Carver c(ws, ws_bytes);
    if (queue_cap(total, total_m) >= 0xFFFFFFFFull)
        return fail(LSTC_ERR_INVALID, "count: more than 2^32-1 work items in one launch");
    uint32_t *ctl = c.take<uint32_t>(64);
    uint32_t *wcnt = c.take<uint32_t>(kPlanEntries);
    unsigned long long *queue = c.take<unsigned long long>(queue_cap(total, total_m));
    if (!ws || !c.ok()) return fail(LSTC_ERR_WORKSPACE, "count: workspace %zu < %zu bytes", ws_bytes, c.need());

    GraphSet gs{};
    gs.G = G;
    gs.total_rows = total;
    if (G == 1) {
        gs.one = to_dev(hosts[0]);
    } else {
        gs.table = d_table;
        gs.row_base = d_row_base;
    }
    for (uint32_t g = 0; g < G; ++g)
        if (hosts[g].per_row) LSTC_TRY(fill_u64(hosts[g].per_row, 0, hosts[g].n_owned, st));
    hipLaunchKernelGGL(k_plan_init, dim3(1), dim3(256), 0, st, ctl, d_counts, G);
    LSTC_LAUNCHED("k_plan_init");
    const int cus = device_cus();
    Bounds bd{{16, 64, tr.small_max, 1024, 2048, 0xFFFFFFFFu}, tr.tiny_max};
    if (total) {
        hipLaunchKernelGGL(k_plan_rows<0>, dim3(kPlanBlocks), dim3(256), 0, st, gs, bd, wcnt, queue);
        LSTC_LAUNCHED("k_plan_rows<0>");
        hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(1024), 0, st, wcnt, ctl, tr.small_group,
                           (uint32_t)cus * tr.grid_per_cu, tr.waves);
        LSTC_LAUNCHED("k_plan_scan");
        hipLaunchKernelGGL(k_plan_rows<1>, dim3(kPlanBlocks), dim3(256), 0, st, gs, bd, wcnt, queue);
        LSTC_LAUNCHED("k_plan_rows<1>");
    }
    if (total) {
        uint32_t grid = (uint32_t)cus * tr.grid_per_cu;
        bool narrow = true;
        for (uint32_t g = 0; g < G; ++g)
            narrow = narrow && hosts[g].m <= (1ull << 30) && hosts[g].scol_len <= (1ull << 30);
        using Kern = void (*)(GraphSet, const unsigned long long *, uint32_t *, unsigned long long *, uint32_t);
        // [unroll 4/8][narrow][waves 2/4/8]
        static const Kern kerns[2][2][3] = {
            {{k_count<4, false, 2>, k_count<4, false, 4>, k_count<4, false, 8>},
             {k_count<4, true, 2>, k_count<4, true, 4>, k_count<4, true, 8>}},
            {{k_count<8, false, 2>, k_count<8, false, 4>, k_count<8, false, 8>},
             {k_count<8, true, 2>, k_count<8, true, 4>, k_count<8, true, 8>}}};
        const Kern kern = kerns[tr.unroll == 8][narrow][tr.waves == 2 ? 0 : tr.waves == 4 ? 1 : 2];
        // a persistent grid beyond what fits at once would only wait for the queue to drain
        const uint32_t fit = resident_blocks((const void *)kern, tr.waves * 64);
        if (fit && grid > (uint32_t)cus * fit) grid = (uint32_t)cus * fit;
        const TimingMark tm = timed ? timing_begin(st) : TimingMark{nullptr, nullptr};
        hipLaunchKernelGGL(kern, dim3(grid), dim3(tr.waves * 64), 0, st, gs, queue, ctl, d_counts, tr.table_scale);
        LSTC_LAUNCHED("k_count");
        timing_end(tm, st);
    }
    return LSTC_OK;
}

}  // namespace lstc
