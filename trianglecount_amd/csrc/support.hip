// support.hip -- per-edge triangle support and per-vertex triangle counts.
//
// The reference keeps per-source partials g_sum[u] (the triangles whose
// lowest-ranked vertex is u) before it reduces them
// (src/WithoutPartition/ShareKernelTCV62_update_v1.cu:175-186); local
// clustering and k-truss need the counts per vertex and per edge instead.
// Every triangle u -> w -> x (u -> x, u owned) is found once, at the arc u -> w,
// as an entry x of N+(w) that N+(u) also holds, and adds one to the support of
// its three arcs and to the count of its three vertices:
//
//   one wave per owned row u (and per tile of kSupTile keys of N+(u), for longer
//   rows): the tile goes into an LDS hash table of 4-slot buckets (key -> its
//   index in the tile), with one counter per key (the arcs u -> x);
//   the lists N+(w), w in N+(u), are walked flat, 64 segments per window and 64
//   entries per round, kSupUnroll rounds with their loads in flight together;
//   a lane's segment is the number of segment starts at or before its position
//   (a ballot of start marks in LDS), and its entry x is looked up in the table;
//   a hit adds to the tile counter of x (LDS), to the window counter of w (LDS),
//   and, with a global atomic, to the support of arc w -> x; window and tile
//   counters go out with one atomic each, to the supports of arcs u -> w and
//   u -> x and to the counts of w and x (so a hub vertex receives one atomic per
//   arc of its, not one per triangle of its).
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint32_t kSupTile = 512;            // keys of N+(u) per tile
static constexpr uint32_t kSupSlots = 2 * kSupTile;  // table slots: load <= 1/2
static constexpr uint32_t kSupBuckets = kSupSlots / 4;
static constexpr int kSupWaves = 4;
static constexpr int kSupUnroll = 8;             // rounds of 64 entries whose loads are in flight together
static constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;  // ids are below 2^32 - 1

__device__ __forceinline__ uint32_t sup_home(uint32_t x) { return (x * 0x9E3779B1u) >> 24; }  // 256 buckets

struct SupWave {
    uint4 tab[kSupBuckets];       // keys, 4 slots per bucket, filled in slot order
    uint16_t tidx[kSupSlots];     // tile index of the key in each slot
    uint32_t cnt[kSupTile];       // triangles found per tile key (arc u -> x)
    uint32_t start[64];           // window: flat start of each non-empty segment
    uint64_t base[64];            //         column index of its first entry
    uint32_t nb[64];              //         its neighbour index in N+(u) (arc u -> w)
    uint32_t hits[64];            //         triangles found per segment
    uint8_t mark[64];             //         segment starts inside the current round
};

// Tile index of x in the table, or -1.  A key lies in the first bucket from its
// home that had a free slot when it was inserted, and slots never empty, so the
// walk ends at the key or at a bucket that is not full; at most half of the
// slots are taken, so that happens within one lap (the bound).
__device__ __forceinline__ int sup_find(const SupWave &S, uint32_t x) {
    uint32_t b = sup_home(x);
    for (uint32_t step = 0; step < kSupBuckets; ++step) {
        const uint4 v = S.tab[b];
        if (v.x == x) return S.tidx[4 * b];
        if (v.y == x) return S.tidx[4 * b + 1];
        if (v.z == x) return S.tidx[4 * b + 2];
        if (v.w == x) return S.tidx[4 * b + 3];
        if (v.w == kEmptySlot) return -1;
        b = (b + 1) & (kSupBuckets - 1);
    }
    return -1;
}

__global__ __launch_bounds__(kSupWaves * 64) void k_support(const uint64_t *__restrict__ off,
                                                            const uint32_t *__restrict__ col, uint32_t n_rows,
                                                            uint32_t n_owned, uint32_t *__restrict__ support,
                                                            unsigned long long *__restrict__ per_vertex) {
    __shared__ SupWave s_w[kSupWaves];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    SupWave &S = s_w[wv];
    uint32_t *slot = (uint32_t *)S.tab;
    const uint64_t le_mask = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const uint64_t nw = (uint64_t)gridDim.x * kSupWaves;
    for (uint64_t u = (uint64_t)blockIdx.x * kSupWaves + wv; u < n_owned; u += nw) {
        const uint64_t rb = off[u];
        const uint32_t K = (uint32_t)(off[u + 1] - rb);
        if (K < 2) continue;
        unsigned long long row_tri = 0;
        for (uint32_t t0 = 0; t0 < K; t0 += kSupTile) {
            const uint32_t T = min(K - t0, kSupTile);
            for (uint32_t i = lane; i < kSupSlots; i += 64) slot[i] = kEmptySlot;
            for (uint32_t i = lane; i < T; i += 64) S.cnt[i] = 0;
            for (uint32_t i = lane; i < T; i += 64) {
                const uint32_t key = col[rb + t0 + i];
                uint32_t b = sup_home(key);
                for (uint32_t step = 0; step < kSupBuckets; ++step) {  // a slot is tried once the ones before it are taken
                    uint32_t *q = slot + 4 * b;
                    int sl = -1;
                    if (atomicCAS(q, kEmptySlot, key) == kEmptySlot) sl = 0;
                    else if (atomicCAS(q + 1, kEmptySlot, key) == kEmptySlot) sl = 1;
                    else if (atomicCAS(q + 2, kEmptySlot, key) == kEmptySlot) sl = 2;
                    else if (atomicCAS(q + 3, kEmptySlot, key) == kEmptySlot) sl = 3;
                    if (sl >= 0) {
                        S.tidx[4 * b + sl] = (uint16_t)i;
                        break;
                    }
                    b = (b + 1) & (kSupBuckets - 1);
                }
            }
            const uint32_t lo = col[rb + t0], hi = col[rb + t0 + T - 1];
            const bool tiled = K > kSupTile;
            for (uint32_t j0 = 0; j0 < K; j0 += 64) {
                const uint32_t j = j0 + lane;
                uint64_t wb = 0;
                uint32_t len = 0;
                if (j < K) {
                    const uint32_t w = col[rb + j];
                    if (w < n_rows) {
                        wb = off[w];
                        len = (uint32_t)(off[w + 1] - wb);
                        // a list that cannot meet this tile is skipped (only when N+(u) is tiled)
                        if (tiled && len && (col[wb + len - 1] < lo || col[wb] > hi)) len = 0;
                    }
                }
                const uint64_t ne = ballot(len > 0);
                if (ne == 0) continue;
                const uint32_t incl = wave_incl_scan(len, lane);
                const uint32_t total = read_lane(incl, 63);
                const uint32_t st = incl - len;  // this lane's segment start (when non-empty)
                if (len > 0) {                   // non-empty segments, compacted in order
                    const uint32_t c = mask_rank(ne);
                    S.start[c] = st;
                    S.base[c] = wb;
                    S.nb[c] = j;
                    S.hits[c] = 0;
                }
                S.mark[lane] = 0;
                for (uint32_t B = 0; B < total; B += 64 * kSupUnroll) {
                    uint64_t pos[kSupUnroll];
                    uint32_t sg[kSupUnroll], x[kSupUnroll];
                    bool ok[kSupUnroll];
#pragma unroll
                    for (int r = 0; r < kSupUnroll; ++r) {
                        const uint32_t Br = B + 64 * r;
                        // segment of position Br + lane: the last one starting at or before it
                        const uint32_t j0r = (uint32_t)__popcll(ballot(len > 0 && st <= Br)) - 1u;
                        const bool in = len > 0 && st > Br && st < Br + 64;
                        if (in) S.mark[st - Br] = 1;
                        const uint64_t marks = ballot(S.mark[lane] != 0);
                        if (in) S.mark[st - Br] = 0;
                        sg[r] = j0r + (uint32_t)__popcll(marks & le_mask);
                        ok[r] = Br + lane < total;
                        pos[r] = ok[r] ? S.base[sg[r]] + (Br + lane - S.start[sg[r]]) : 0u;
                    }
#pragma unroll
                    for (int r = 0; r < kSupUnroll; ++r) x[r] = ok[r] ? col[pos[r]] : kEmptySlot;
#pragma unroll
                    for (int r = 0; r < kSupUnroll; ++r) {
                        if (x[r] < lo || x[r] > hi) continue;  // also every position past the end
                        const int i = sup_find(S, x[r]);
                        if (i >= 0) {
                            atomicAdd(&S.cnt[i], 1u);                       // arc u -> x
                            atomicAdd(&S.hits[sg[r]], 1u);                  // arc u -> w
                            if (support) atomicAdd(&support[pos[r]], 1u);  // arc w -> x
                        }
                    }
                }
                if (lane < (uint32_t)__popcll(ne)) {
                    const uint32_t h = S.hits[lane];
                    if (h) {
                        const uint64_t a = rb + S.nb[lane];
                        if (support) atomicAdd(&support[a], h);
                        if (per_vertex) atomicAdd(&per_vertex[col[a]], (unsigned long long)h);
                        row_tri += h;
                    }
                }
            }
            for (uint32_t i = lane; i < T; i += 64) {
                const uint32_t c = S.cnt[i];
                if (!c) continue;
                if (support) atomicAdd(&support[rb + t0 + i], c);                                // arc u -> x
                if (per_vertex) atomicAdd(&per_vertex[col[rb + t0 + i]], (unsigned long long)c);  // vertex x
            }
        }
        const unsigned long long tot = wave_sum64(row_tri);
        if (per_vertex && lane == 0 && tot) atomicAdd(&per_vertex[u], tot);
    }
}

// ---------------------------------------------------------------- split rule
// With target-side in-lists the same outputs come from the split rule of the
// count (count_engine.hip): arc u -> w is handled at u (kind A: N+(u) in the
// table, N+(w) streamed) when d+(w) < d+(u), else at w (kind B: N+(w) in the
// table, N+(u) streamed for u in w's in-list), so every arc streams its shorter
// list (at RMAT s23 2.9e10 entries instead of 7.3e10).  A triangle u -> w -> x
// found at row r adds to the tile counter of x (arc r -> x), to the segment
// counter of the streamed list (arc r -> w for kind A; arc u -> r for kind B,
// whose position in N+(u) one binary search finds per segment with hits), and
// with a global atomic to the streamed entry (arc w -> x, or u -> x).
// Work items: chunk 0 of a row holds its kind A side, chunk c its in-list
// sources [512c, 512c + 512), so a hub's long in-list is spread over waves;
// persistent waves take items from a ticket.
static constexpr uint32_t kSupChunk = 512;

__global__ void k_sup_items(const uint64_t *__restrict__ off, const uint64_t *__restrict__ in_off, uint32_t n_rows,
                            uint32_t n_owned, uint32_t *__restrict__ items) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t K = off[r + 1] - off[r], kb = in_off[r + 1] - in_off[r];
        const uint32_t nb = (uint32_t)((kb + kSupChunk - 1) / kSupChunk);
        const uint32_t a = (r < n_owned && K >= 2) ? 1u : 0u;
        items[r] = K == 0 ? 0u : (nb > a ? nb : a);
    }
}

__global__ void k_sup_item_rows(const uint32_t *__restrict__ items, const uint64_t *__restrict__ item_off,
                                uint32_t n_rows, uint32_t *__restrict__ item_row) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows; r += (uint64_t)gridDim.x * blockDim.x)
        for (uint32_t c = 0; c < items[r]; ++c) item_row[item_off[r] + c] = (uint32_t)r;
}

// Tile index of x, the first bucket read inline; the walk on is rare (a full home).
__device__ __forceinline__ int sup_lookup(const SupWave &S, uint32_t x) {
    const uint32_t b = sup_home(x);
    const uint4 v = S.tab[b];
    const int sl = v.x == x ? 0 : v.y == x ? 1 : v.z == x ? 2 : v.w == x ? 3 : -1;
    if (sl >= 0) return S.tidx[4 * b + sl];
    if (v.w == kEmptySlot) return -1;
    return sup_find(S, x);
}

// Position of key r in the sorted list col[b, e) (r is there).
__device__ __forceinline__ uint64_t sup_position(const uint32_t *__restrict__ col, uint64_t b, uint64_t e, uint32_t r) {
    while (e - b > 1) {
        const uint64_t mid = (b + e) >> 1;
        if (col[mid] <= r) b = mid; else e = mid;
    }
    return b;
}

__global__ __launch_bounds__(kSupWaves * 64) void k_support_split(
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ col, const uint64_t *__restrict__ in_off,
    const uint32_t *__restrict__ in_col, uint32_t n_rows, uint32_t n_owned, const uint64_t *__restrict__ item_off,
    const uint32_t *__restrict__ item_row, uint32_t *__restrict__ ticket, uint32_t *__restrict__ support,
    unsigned long long *__restrict__ per_vertex) {
    __shared__ SupWave s_w[kSupWaves];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    SupWave &S = s_w[wv];
    uint32_t *slot = (uint32_t *)S.tab;
    const uint64_t le_mask = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const uint64_t total = item_off[n_rows];
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(ticket, 1u);
        t = read_lane(t, 0);
        if (t >= total) break;  // every wave reaches this once the items are taken
        const uint32_t r = item_row[t];
        const uint32_t chunk = (uint32_t)(t - item_off[r]);
        const uint64_t rb = off[r];
        const uint32_t K = (uint32_t)(off[r + 1] - rb);
        const uint64_t ib = in_off[r], kb = in_off[r + 1] - ib;
        const uint32_t Ka = (chunk == 0 && r < n_owned && K >= 2) ? K : 0u;
        const uint64_t b0 = (uint64_t)chunk * kSupChunk;
        const uint32_t Kb = b0 < kb ? (uint32_t)min(kb - b0, (uint64_t)kSupChunk) : 0u;
        const uint32_t C = Ka + Kb;  // candidates: kind A neighbours, then kind B sources
        if (C == 0) continue;
        unsigned long long row_tri = 0;
        for (uint32_t t0 = 0; t0 < K; t0 += kSupTile) {
            const uint32_t T = min(K - t0, kSupTile);
            for (uint32_t i = lane; i < kSupSlots; i += 64) slot[i] = kEmptySlot;
            for (uint32_t i = lane; i < T; i += 64) S.cnt[i] = 0;
            for (uint32_t i = lane; i < T; i += 64) {
                const uint32_t key = col[rb + t0 + i];
                uint32_t b = sup_home(key);
                for (uint32_t step = 0; step < kSupBuckets; ++step) {  // a slot is tried once the ones before it are taken
                    uint32_t *q = slot + 4 * b;
                    int sl = -1;
                    if (atomicCAS(q, kEmptySlot, key) == kEmptySlot) sl = 0;
                    else if (atomicCAS(q + 1, kEmptySlot, key) == kEmptySlot) sl = 1;
                    else if (atomicCAS(q + 2, kEmptySlot, key) == kEmptySlot) sl = 2;
                    else if (atomicCAS(q + 3, kEmptySlot, key) == kEmptySlot) sl = 3;
                    if (sl >= 0) {
                        S.tidx[4 * b + sl] = (uint16_t)i;
                        break;
                    }
                    b = (b + 1) & (kSupBuckets - 1);
                }
            }
            const uint32_t lo = col[rb + t0], hi = col[rb + t0 + T - 1];
            const bool tiled = K > kSupTile;
            for (uint32_t c0 = 0; c0 < C; c0 += 64) {
                const uint32_t c = c0 + lane;
                uint64_t wb = 0;
                uint32_t len = 0;
                if (c < Ka) {  // kind A: N+(w) when it is the shorter list
                    const uint32_t w = col[rb + c];
                    if (w < n_rows) {
                        wb = off[w];
                        len = (uint32_t)(off[w + 1] - wb);
                        if (len >= K) len = 0;
                    }
                } else if (c < C) {  // kind B: N+(u) of an in-list source u (d+(u) <= d+(r))
                    const uint32_t u = in_col[ib + b0 + (c - Ka)];
                    wb = off[u];
                    len = (uint32_t)(off[u + 1] - wb);
                }
                // a list that cannot meet this tile is skipped (only when N+(r) is tiled)
                if (tiled && len && (col[wb + len - 1] < lo || col[wb] > hi)) len = 0;
                const uint64_t ne = ballot(len > 0);
                if (ne == 0) continue;
                const uint32_t incl = wave_incl_scan(len, lane);
                const uint32_t total_e = read_lane(incl, 63);
                const uint32_t st = incl - len;
                if (len > 0) {
                    const uint32_t k = mask_rank(ne);
                    S.start[k] = st;
                    S.base[k] = wb;
                    S.nb[k] = c;
                    S.hits[k] = 0;
                }
                S.mark[lane] = 0;
                for (uint32_t B = 0; B < total_e; B += 64 * kSupUnroll) {
                    uint64_t pos[kSupUnroll];
                    uint32_t sg[kSupUnroll], x[kSupUnroll];
                    bool ok[kSupUnroll];
#pragma unroll
                    for (int q = 0; q < kSupUnroll; ++q) {
                        const uint32_t Br = B + 64 * q;
                        const uint32_t j0r = (uint32_t)__popcll(ballot(len > 0 && st <= Br)) - 1u;
                        const bool in = len > 0 && st > Br && st < Br + 64;
                        if (in) S.mark[st - Br] = 1;
                        const uint64_t marks = ballot(S.mark[lane] != 0);
                        if (in) S.mark[st - Br] = 0;
                        sg[q] = j0r + (uint32_t)__popcll(marks & le_mask);
                        ok[q] = Br + lane < total_e;
                        pos[q] = ok[q] ? S.base[sg[q]] + (Br + lane - S.start[sg[q]]) : 0u;
                    }
#pragma unroll
                    for (int q = 0; q < kSupUnroll; ++q) x[q] = ok[q] ? col[pos[q]] : kEmptySlot;
#pragma unroll
                    for (int q = 0; q < kSupUnroll; ++q) {
                        if (x[q] < lo || x[q] > hi) continue;  // also every position past the end
                        const int i = sup_lookup(S, x[q]);
                        if (i >= 0) {
                            atomicAdd(&S.cnt[i], 1u);                       // arc r -> x
                            atomicAdd(&S.hits[sg[q]], 1u);                  // arc r -> w, or u -> r
                            if (support) atomicAdd(&support[pos[q]], 1u);  // arc w -> x, or u -> x
                        }
                    }
                }
                if (lane < (uint32_t)__popcll(ne)) {
                    const uint32_t h = S.hits[lane];
                    if (h) {
                        const uint32_t cc = S.nb[lane];
                        if (cc < Ka) {
                            const uint64_t a = rb + cc;
                            if (support) atomicAdd(&support[a], h);
                            if (per_vertex) atomicAdd(&per_vertex[col[a]], (unsigned long long)h);
                        } else {
                            const uint32_t u = in_col[ib + b0 + (cc - Ka)];
                            if (support) atomicAdd(&support[sup_position(col, off[u], off[u + 1], r)], h);
                            if (per_vertex) atomicAdd(&per_vertex[u], (unsigned long long)h);
                        }
                        row_tri += h;
                    }
                }
            }
            for (uint32_t i = lane; i < T; i += 64) {
                const uint32_t c = S.cnt[i];
                if (!c) continue;
                if (support) atomicAdd(&support[rb + t0 + i], c);                                // arc r -> x
                if (per_vertex) atomicAdd(&per_vertex[col[rb + t0 + i]], (unsigned long long)c);  // vertex x
            }
        }
        const unsigned long long tot = wave_sum64(row_tri);
        if (per_vertex && lane == 0 && tot) atomicAdd(&per_vertex[r], tot);
    }
}

struct SupLayout {
    uint32_t *items, *item_row, *ticket;
    uint64_t *item_off;
    void *scan_ws;
    size_t scan_bytes, bytes;
};

static SupLayout sup_layout(void *ws, size_t ws_bytes, uint32_t n_rows, uint64_t m) {
    Carver c(ws, ws_bytes);
    SupLayout L;
    L.items = c.take<uint32_t>((size_t)n_rows + 1);
    L.item_off = c.take<uint64_t>((size_t)n_rows + 1);
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
    if (d_support && m) LSTC_TRY(fill_u32(d_support, 0, m, st));
    if (d_per_vertex && n_rows) LSTC_TRY(fill_u64(d_per_vertex, 0, n_rows, st));
    if (n_owned == 0 || m == 0) return LSTC_OK;
    const uint32_t grid = (uint32_t)hmin<uint64_t>(((uint64_t)n_owned + kSupWaves - 1) / kSupWaves,
                                                   (uint64_t)device_cus() * 4);
    hipLaunchKernelGGL(k_support, dim3(grid), dim3(kSupWaves * 64), 0, st, d_off, d_col, n_rows, n_owned, d_support,
                       (unsigned long long *)d_per_vertex);
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
    if (2 * (uint64_t)n_rows + m / kSupChunk + 1 >= 0xFFFFFFFFull)
        return fail(LSTC_ERR_INVALID, "lstc_triangle_support_split: too many work items");
    const SupLayout L = sup_layout(d_ws, ws_bytes, n_rows, m);
    if (!d_ws || ws_bytes < L.bytes)
        return fail(LSTC_ERR_WORKSPACE, "lstc_triangle_support_split: workspace %zu < %zu", ws_bytes, L.bytes);
    if (d_support && m) LSTC_TRY(fill_u32(d_support, 0, m, st));
    if (d_per_vertex && n_rows) LSTC_TRY(fill_u64(d_per_vertex, 0, n_rows, st));
    if (n_rows == 0 || m == 0) return LSTC_OK;
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n_rows + 255) / 256, (uint64_t)device_cus() * 16);
    hipLaunchKernelGGL(k_sup_items, dim3(grid_n), dim3(256), 0, st, d_off, d_in_off, n_rows, n_owned, L.items);
    LSTC_LAUNCHED("k_sup_items");
    LSTC_TRY(scan_exclusive_u32_to_u64(L.items, L.item_off, n_rows, L.item_off + n_rows, L.scan_ws, L.scan_bytes, st));
    hipLaunchKernelGGL(k_sup_item_rows, dim3(grid_n), dim3(256), 0, st, L.items, L.item_off, n_rows, L.item_row);
    LSTC_LAUNCHED("k_sup_item_rows");
    LSTC_TRY(fill_u32(L.ticket, 0, 1, st));
    const uint32_t grid = (uint32_t)device_cus() * 4;  // persistent: four workgroups of four waves per CU (LDS)
    hipLaunchKernelGGL(k_support_split, dim3(grid), dim3(kSupWaves * 64), 0, st, d_off, d_col, d_in_off, d_in_col,
                       n_rows, n_owned, (const uint64_t *)L.item_off, (const uint32_t *)L.item_row, L.ticket, d_support,
                       (unsigned long long *)d_per_vertex);
    LSTC_LAUNCHED("k_support_split");
    return LSTC_OK;
}
