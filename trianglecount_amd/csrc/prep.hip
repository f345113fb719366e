// prep.hip -- (a) GPU clean + orient into a uint32 CSR with uint64 offsets.
//
// Replaces the DGL cleaning of Med_dataset/without_part_preprocess.py:193-208
// (remove_self_loop, to_bidirected, isolated-node removal) and the CuPy
// orientation Find_size / cp.cumsum / Convert of :19-111,253-270.
//
// Pipeline (all stream-ordered, sizes kept on the device until the end):
//   keys  : each raw pair (s, d), s != d, emits (s<<b | d) and (d<<b | s);
//           loops and out-of-range ids emit a sentinel that sorts last
//   sort  : LSD radix sort over 2b bits (prims.hip)
//   then three compactions of the sorted list, each a count pass (per tile of
//   4096 keys) + a scan of the tile counts + a scatter pass that recomputes the
//   predicate and writes the kept entries at their final positions (no flag or
//   position arrays); the scatter also records each row's first / end position:
//   unique: first of each run          -> symmetric arc list, row bounds, deg[v]
//   relabel (compact): id' = rank of v among vertices with deg > 0 (monotone)
//   symmetric CSR: offsets = scan(deg) compacted, columns relabelled
//   orient: keep s->d iff deg s < deg d or (equal and d < s)
//           -> oriented columns (rows ascending), d+(v), offsets = scan(d+)
//   in    : reverse keys of kept arcs with d+(d) <= d+(s) -> target-side in-lists
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint64_t kSentinel = ~0ull;

static int bits_for(uint64_t n) {  // bits to represent values < n
    int b = 0;
    while (b < 64 && ((n - 1) >> b)) ++b;
    return b == 0 ? 1 : b;
}

#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

__global__ void k_make_keys(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                            uint64_t m_raw, uint32_t n, uint32_t b, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ bad) {
    GRID_STRIDE(i, m_raw) {
        const uint32_t s = src[i], d = dst[i];
        uint64_t k0 = kSentinel, k1 = kSentinel;
        if (s >= n || d >= n) {
            *bad = 1;
        } else if (s != d) {
            k0 = ((uint64_t)s << b) | d;
            k1 = ((uint64_t)d << b) | s;
        }
        keys[2 * i] = k0;
        keys[2 * i + 1] = k1;
    }
}

// ------------------------------------------------------------------ compactions
// Tile = 256 threads x 16 rounds = 4096 keys; key (round r, thread t) is tile
// position r * 256 + t, so loads are coalesced and the order inside the tile is
// (round, wave, lane).  The scatter pass ranks the kept keys of a tile in that
// order from one ballot per round and a 64-entry prefix over (round, wave).
static constexpr uint32_t kCT = 256;
static constexpr uint32_t kCR = 16;
static constexpr uint64_t kCTile = (uint64_t)kCT * kCR;

template <class Op>
__global__ __launch_bounds__(kCT) void k_compact_count(Op op, uint32_t *__restrict__ tile_cnt) {
    __shared__ uint32_t s_w[kCT / 64];
    const uint64_t n = op.size();
    const uint64_t tile0 = (uint64_t)blockIdx.x * kCTile;
    uint32_t c = 0;
    if (tile0 < n) {
#pragma unroll 4
        for (uint32_t r = 0; r < kCR; ++r) {
            const uint64_t i = tile0 + (uint64_t)r * kCT + threadIdx.x;
            if (i < n && op.pred(i)) ++c;
        }
    }
    c = wave_sum32(c);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

template <class Op>
__global__ __launch_bounds__(kCT) void k_compact_scatter(Op op, const uint64_t *__restrict__ tile_off) {
    __shared__ uint32_t s_pre[kCR * (kCT / 64)];  // kept keys per (round, wave), then their exclusive prefix
    const uint64_t n = op.size();
    const uint64_t tile0 = (uint64_t)blockIdx.x * kCTile;
    if (tile0 >= n) return;  // the whole block leaves together
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t kept[kCR];
#pragma unroll
    for (uint32_t r = 0; r < kCR; ++r) {
        const uint64_t i = tile0 + (uint64_t)r * kCT + threadIdx.x;
        kept[r] = ballot(i < n && op.pred(i));
        if (lane == 0) s_pre[r * (kCT / 64) + wv] = (uint32_t)__popcll(kept[r]);
    }
    __syncthreads();
    if (wv == 0) {
        const uint32_t v = s_pre[lane];
        const uint32_t incl = wave_incl_scan(v, lane);
        s_pre[lane] = incl - v;
    }
    __syncthreads();
    const uint64_t base = tile_off[blockIdx.x];
#pragma unroll
    for (uint32_t r = 0; r < kCR; ++r) {
        const uint64_t i = tile0 + (uint64_t)r * kCT + threadIdx.x;
        if (i < n) op.emit(i, lane_in(kept[r]), base + s_pre[r * (kCT / 64) + wv] + mask_rank(kept[r]));
    }
}

// The row of key i starts (ends) there when the previous (next) key has another row.
__device__ __forceinline__ bool row_first(const uint64_t *k, uint64_t i, uint32_t b, uint64_t s) {
    return i == 0 || (k[i - 1] >> b) != s;
}
__device__ __forceinline__ bool row_last(const uint64_t *k, uint64_t i, uint64_t n, uint32_t b, uint64_t s) {
    return i + 1 == n || (k[i + 1] >> b) != s;
}

// Unique keys of the sorted list (sentinels dropped), with each row's bounds
// [first, last) in the unique list.
struct UniqueOp {
    const uint64_t *keys;
    uint64_t n;
    uint32_t b;
    uint64_t *uk, *first, *last;
    __device__ uint64_t size() const { return n; }
    __device__ bool pred(uint64_t i) const {
        const uint64_t k = keys[i];
        return k != kSentinel && (i == 0 || keys[i - 1] != k);
    }
    __device__ void emit(uint64_t i, bool kept, uint64_t pos) const {
        const uint64_t k = keys[i];
        if (k == kSentinel) return;
        if (kept) uk[pos] = k;
        const uint64_t s = k >> b;
        if (row_first(keys, i, b, s)) first[s] = pos;  // a row's first key is always kept
        if (row_last(keys, i, n, b, s)) last[s] = pos + (kept ? 1u : 0u);
    }
};

// Oriented arcs (keep s -> d iff deg s < deg d, or equal and d < s), written as
// relabelled ids; end[s] - start[s] = d+(s).
struct OrientOp {
    const uint64_t *uk;
    const uint64_t *d_m2;
    uint32_t b;
    uint64_t mask;
    const uint2 *dd;  // {deg, d+} per vertex
    const uint32_t *nid;
    uint32_t *col;
    uint64_t *start, *end;
    __device__ uint64_t size() const { return *d_m2; }
    __device__ bool pred(uint64_t i) const {
        const uint64_t k = uk[i];
        const uint32_t s = (uint32_t)(k >> b), d = (uint32_t)(k & mask);
        const uint32_t ds = dd[s].x, dg = dd[d].x;
        return ds < dg || (ds == dg && d < s);
    }
    __device__ void emit(uint64_t i, bool kept, uint64_t pos) const {
        const uint64_t k = uk[i], s = k >> b;
        if (kept) col[pos] = nid[k & mask];
        const uint64_t n = *d_m2;
        if (row_first(uk, i, b, s)) start[s] = pos;
        if (row_last(uk, i, n, b, s)) end[s] = pos + (kept ? 1u : 0u);
    }
};

// Target-side in-lists: in row s, the sources d of arcs d -> s (the reverse of
// an oriented arc) with d+(d) <= d+(s): the arcs the split count handles from
// the target side.  Rows stay ascending.
struct InOp {
    const uint64_t *uk;
    const uint64_t *d_m2;
    uint32_t b;
    uint64_t mask;
    const uint2 *dd;
    const uint32_t *nid;
    uint32_t *col;
    uint64_t *start, *end;
    __device__ uint64_t size() const { return *d_m2; }
    __device__ bool pred(uint64_t i) const {
        const uint64_t k = uk[i];
        const uint32_t s = (uint32_t)(k >> b), d = (uint32_t)(k & mask);
        const uint2 vs = dd[s], vd = dd[d];  // one 8-byte gather per endpoint
        const bool out_arc = vs.x < vd.x || (vs.x == vd.x && d < s);
        return !out_arc && vd.y <= vs.y;
    }
    __device__ void emit(uint64_t i, bool kept, uint64_t pos) const {
        const uint64_t k = uk[i], s = k >> b;
        if (kept) col[pos] = nid[k & mask];
        const uint64_t n = *d_m2;
        if (row_first(uk, i, b, s)) start[s] = pos;
        if (row_last(uk, i, n, b, s)) end[s] = pos + (kept ? 1u : 0u);
    }
};

template <class Op>
static lstc_status compact_pass(const Op &op, uint64_t n_max, uint32_t *tile_cnt, uint64_t *tile_off, uint64_t *d_total,
                           void *scan_ws, size_t scan_bytes, hipStream_t st) {
    const uint64_t tiles = (n_max + kCTile - 1) / kCTile;
    if (tiles == 0) return fill_u64(d_total, 0, 1, st);
    hipLaunchKernelGGL(k_compact_count<Op>, dim3((uint32_t)tiles), dim3(kCT), 0, st, op, tile_cnt);
    LSTC_LAUNCHED("k_compact_count");
    LSTC_TRY(scan_exclusive_u32_to_u64(tile_cnt, tile_off, tiles, d_total, scan_ws, scan_bytes, st));
    hipLaunchKernelGGL(k_compact_scatter<Op>, dim3((uint32_t)tiles), dim3(kCT), 0, st, op, (const uint64_t *)tile_off);
    LSTC_LAUNCHED("k_compact_scatter");
    return LSTC_OK;
}

// Row lengths from the compactions' row bounds; deg also goes into dd[v].x, d+ into dd[v].y.
__global__ void k_row_lengths(const uint64_t *__restrict__ first, const uint64_t *__restrict__ last, uint32_t n,
                              uint32_t *__restrict__ len, uint2 *__restrict__ dd, int field) {
    GRID_STRIDE(v, n) {
        const uint32_t l = (uint32_t)(last[v] - first[v]);
        len[v] = l;
        if (dd) {
            if (field == 0) dd[v] = make_uint2(l, 0u);
            else dd[v].y = l;
        }
    }
}

__global__ void k_nonzero(const uint32_t *__restrict__ deg, uint32_t n, int compact,
                          uint32_t *__restrict__ f) {
    GRID_STRIDE(v, n) f[v] = (!compact || deg[v] != 0) ? 1u : 0u;
}

// nid[v] = new id (from the exclusive scan of the keep flags); writes compacted offsets.
__global__ void k_relabel_offsets(const uint32_t *__restrict__ keepv, const uint64_t *__restrict__ nidscan,
                                  const uint64_t *__restrict__ degscan, uint32_t n,
                                  const uint64_t *__restrict__ d_n_out, const uint64_t *__restrict__ d_m2,
                                  uint32_t *__restrict__ nid, uint64_t *__restrict__ sym_off) {
    GRID_STRIDE(v, n) {
        if (keepv[v]) {
            const uint32_t id = (uint32_t)nidscan[v];
            nid[v] = id;
            if (sym_off) sym_off[id] = degscan[v];
        } else {
            nid[v] = 0xFFFFFFFFu;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && sym_off) sym_off[*d_n_out] = *d_m2;
}

__global__ void k_sym_cols(const uint64_t *__restrict__ uk, const uint64_t *__restrict__ d_m2,
                           uint64_t mask, const uint32_t *__restrict__ nid, uint32_t *__restrict__ sym_col) {
    const uint64_t m2 = *d_m2;
    GRID_STRIDE(i, m2) sym_col[i] = nid[uk[i] & mask];
}

__global__ void k_orient_offsets(const uint32_t *__restrict__ keepv, const uint64_t *__restrict__ nidscan,
                                 const uint64_t *__restrict__ doscan, uint32_t n,
                                 const uint64_t *__restrict__ d_n_out, const uint64_t *__restrict__ d_m,
                                 uint64_t *__restrict__ off) {
    GRID_STRIDE(v, n) if (keepv[v]) off[nidscan[v]] = doscan[v];
    if (blockIdx.x == 0 && threadIdx.x == 0) off[*d_n_out] = *d_m;
}

// Segment words (start | length << 32 of the row of every entry) of the
// oriented columns and the in-lists, with the sizes still on the device:
// sizes[1] rows, sizes[2] arcs, sizes[3] in-list entries.
__global__ void k_prep_seg_words(const uint64_t *__restrict__ sizes, const uint64_t *__restrict__ off,
                                 const uint32_t *__restrict__ col, const uint32_t *__restrict__ in_col,
                                 uint64_t *__restrict__ col_seg, uint64_t *__restrict__ in_seg) {
    const uint64_t n_rows = sizes[1], m = sizes[2], n_in = sizes[3];
    GRID_STRIDE(i, m + n_in) {
        const bool in = i >= m;
        const uint32_t v = in ? in_col[i - m] : col[i];
        uint64_t w = 0;
        if (v < n_rows) {
            const uint64_t b = off[v];
            w = (uint64_t)(uint32_t)b | ((off[v + 1] - b) << 32);
        }
        if (in) in_seg[i - m] = w; else col_seg[i] = w;
    }
}

struct PrepLayout {
    uint64_t *keys, *alt;
    uint32_t *tile_cnt;
    uint64_t *tile_off;
    uint32_t *deg, *keepv, *nid, *dout, *din;
    uint2 *dd;
    uint64_t *degscan, *nidscan, *doscan, *first, *last, *discan;
    uint64_t *scal;  // [0]=m2 [1]=n_out [2]=m [3]=in-list entries
    uint32_t *bad;
    void *sort_ws;
    size_t sort_bytes;
    void *scan_ws;
    size_t scan_bytes;
    size_t bytes;
};

static PrepLayout prep_layout(void *ws, size_t ws_bytes, uint32_t n, uint64_t m_raw) {
    Carver c(ws, ws_bytes);
    PrepLayout L;
    const uint64_t nk = 2 * m_raw;
    const uint64_t tiles = (nk + kCTile - 1) / kCTile + 1;
    L.keys = c.take<uint64_t>(nk + 1);
    L.alt = c.take<uint64_t>(nk + 1);
    L.tile_cnt = c.take<uint32_t>(tiles);
    L.tile_off = c.take<uint64_t>(tiles);
    L.deg = c.take<uint32_t>((size_t)n + 1);
    L.keepv = c.take<uint32_t>((size_t)n + 1);
    L.nid = c.take<uint32_t>((size_t)n + 1);
    L.dout = c.take<uint32_t>((size_t)n + 1);
    L.din = c.take<uint32_t>((size_t)n + 1);
    L.dd = c.take<uint2>((size_t)n + 1);
    L.discan = c.take<uint64_t>((size_t)n + 1);
    L.degscan = c.take<uint64_t>((size_t)n + 1);
    L.nidscan = c.take<uint64_t>((size_t)n + 1);
    L.doscan = c.take<uint64_t>((size_t)n + 1);
    L.first = c.take<uint64_t>((size_t)n + 1);
    L.last = c.take<uint64_t>((size_t)n + 1);
    L.scal = c.take<uint64_t>(8);
    L.bad = c.take<uint32_t>(4);
    L.sort_bytes = radix_sort_ws_bytes(nk);
    L.sort_ws = c.take<char>(L.sort_bytes);
    L.scan_bytes = scan_ws_bytes(tiles > (uint64_t)n + 1 ? tiles : (uint64_t)n + 1);
    L.scan_ws = c.take<char>(L.scan_bytes);
    L.bytes = c.need();
    return L;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_clean_orient_workspace(uint32_t n, uint64_t m_raw, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_clean_orient_workspace: NULL bytes");
    *bytes = prep_layout(nullptr, 0, n, m_raw).bytes;
    return LSTC_OK;
}

// Enqueues the whole pipeline without synchronising.  On return *d_sizes points
// at device words {m2, n_out, m, -} and *d_bad at the out-of-range flag; both are
// valid once the stream reaches this point.
lstc_status lstc::clean_orient_launch(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw, uint32_t n,
                                      int compact, uint64_t *d_sym_off, uint32_t *d_sym_col, uint64_t *d_off,
                                      uint32_t *d_col, uint64_t *d_in_off, uint32_t *d_in_col, void *d_ws,
                                      size_t ws_bytes, hipStream_t st,
                                      const uint64_t **d_sizes, const uint32_t **d_bad) {
    if (n == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: n must be < 2^32-1");
    if (!d_off) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL output");
    if ((d_sym_off == nullptr) != (d_sym_col == nullptr))
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient: give both symmetric outputs or neither");
    if (m_raw && (!d_src || !d_dst || !d_col)) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL edge arrays");
    PrepLayout L = prep_layout(d_ws, ws_bytes, n, m_raw);
    if (!d_ws || ws_bytes < L.bytes)
        return fail(LSTC_ERR_WORKSPACE, "lstc_clean_orient: workspace %zu < %zu bytes", ws_bytes, L.bytes);
    const uint32_t b = (uint32_t)bits_for(n ? n : 1);
    const uint64_t mask = (b >= 64) ? ~0ull : ((1ull << b) - 1);
    const uint64_t nk = 2 * m_raw;
    const uint32_t grid_k = (uint32_t)hmin<uint64_t>((nk + 255) / 256 + 1, (uint64_t)device_cus() * 32);
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n + 255) / 256 + 1, (uint64_t)device_cus() * 32);

    LSTC_TRY(fill_u64(L.scal, 0, 8, st));
    LSTC_TRY(fill_u32(L.bad, 0, 4, st));
    LSTC_TRY(fill_u32(L.deg, 0, (uint64_t)n + 1, st));
    LSTC_TRY(fill_u32(L.dout, 0, (uint64_t)n + 1, st));
    if (m_raw) {
        hipLaunchKernelGGL(k_make_keys, dim3(grid_k), dim3(256), 0, st, d_src, d_dst, m_raw, n, b, L.keys, L.bad);
        LSTC_LAUNCHED("k_make_keys");
        uint64_t *sorted = nullptr;
        LSTC_TRY(radix_sort_u64(L.keys, L.alt, nk, (int)(2 * b > 64 ? 64 : 2 * b), &sorted, L.sort_ws, L.sort_bytes, st));
        uint64_t *uk = (sorted == L.keys) ? L.alt : L.keys;
        // unique arcs and the symmetric row bounds (m2 -> scal[0])
        LSTC_TRY(fill_u64(L.first, 0, (uint64_t)n + 1, st));
        LSTC_TRY(fill_u64(L.last, 0, (uint64_t)n + 1, st));
        LSTC_TRY(compact_pass(UniqueOp{sorted, nk, b, uk, L.first, L.last}, nk, L.tile_cnt, L.tile_off, L.scal + 0,
                         L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_row_lengths, dim3(grid_n), dim3(256), 0, st, L.first, L.last, n, L.deg, L.dd, 0);
        LSTC_LAUNCHED("k_row_lengths(deg)");
        hipLaunchKernelGGL(k_nonzero, dim3(grid_n), dim3(256), 0, st, L.deg, n, compact, L.keepv);
        LSTC_LAUNCHED("k_nonzero");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.keepv, L.nidscan, n, L.scal + 1, L.scan_ws, L.scan_bytes, st));
        LSTC_TRY(scan_exclusive_u32_to_u64(L.deg, L.degscan, n, nullptr, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_relabel_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.degscan, n,
                           L.scal + 1, L.scal + 0, L.nid, d_sym_off);
        LSTC_LAUNCHED("k_relabel_offsets");
        if (d_sym_col) {
            hipLaunchKernelGGL(k_sym_cols, dim3(grid_k), dim3(256), 0, st, uk, L.scal + 0, mask, L.nid, d_sym_col);
            LSTC_LAUNCHED("k_sym_cols");
        }
        // oriented arcs (m -> scal[2]) and d+ per vertex
        LSTC_TRY(fill_u64(L.first, 0, (uint64_t)n + 1, st));
        LSTC_TRY(fill_u64(L.last, 0, (uint64_t)n + 1, st));
        LSTC_TRY(compact_pass(OrientOp{uk, L.scal + 0, b, mask, L.dd, L.nid, d_col, L.first, L.last}, nk, L.tile_cnt,
                         L.tile_off, L.scal + 2, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_row_lengths, dim3(grid_n), dim3(256), 0, st, L.first, L.last, n, L.dout, L.dd, 1);
        LSTC_LAUNCHED("k_row_lengths(d+)");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.dout, L.doscan, n, nullptr, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                           L.scal + 1, L.scal + 2, d_off);
        LSTC_LAUNCHED("k_orient_offsets");
        if (d_in_off) {
            // in-lists of the arcs counted from the target side (entries -> scal[3])
            LSTC_TRY(fill_u64(L.first, 0, (uint64_t)n + 1, st));
            LSTC_TRY(fill_u64(L.last, 0, (uint64_t)n + 1, st));
            LSTC_TRY(compact_pass(InOp{uk, L.scal + 0, b, mask, L.dd, L.nid, d_in_col, L.first, L.last}, nk, L.tile_cnt,
                             L.tile_off, L.scal + 3, L.scan_ws, L.scan_bytes, st));
            hipLaunchKernelGGL(k_row_lengths, dim3(grid_n), dim3(256), 0, st, L.first, L.last, n, L.din,
                               (uint2 *)nullptr, 0);
            LSTC_LAUNCHED("k_row_lengths(in)");
            LSTC_TRY(scan_exclusive_u32_to_u64(L.din, L.discan, n, nullptr, L.scan_ws, L.scan_bytes, st));
            hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.discan, n,
                               L.scal + 1, L.scal + 3, d_in_off);
            LSTC_LAUNCHED("k_orient_offsets(in)");
        }
    } else {
        // no edges: every vertex is isolated
        hipLaunchKernelGGL(k_nonzero, dim3(grid_n), dim3(256), 0, st, L.deg, n, compact, L.keepv);
        LSTC_LAUNCHED("k_nonzero");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.keepv, L.nidscan, n, L.scal + 1, L.scan_ws, L.scan_bytes, st));
        LSTC_TRY(fill_u64(L.doscan, 0, (uint64_t)n + 1, st));
        hipLaunchKernelGGL(k_relabel_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                           L.scal + 1, L.scal + 0, L.nid, d_sym_off);
        LSTC_LAUNCHED("k_relabel_offsets");
        hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                           L.scal + 1, L.scal + 2, d_off);
        LSTC_LAUNCHED("k_orient_offsets");
        if (d_in_off) {
            hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                               L.scal + 1, L.scal + 3, d_in_off);
            LSTC_LAUNCHED("k_orient_offsets(in)");
        }
    }
    if (d_sizes) *d_sizes = L.scal;
    if (d_bad) *d_bad = L.bad;
    return LSTC_OK;
}

extern "C" lstc_status lstc_clean_orient_split(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                                               uint32_t n, int compact, uint64_t *d_sym_off, uint32_t *d_sym_col,
                                               uint64_t *d_off, uint32_t *d_col, uint64_t *d_in_off,
                                               uint32_t *d_in_col, uint64_t *d_col_seg, uint64_t *d_in_seg,
                                               uint32_t *n_out, uint64_t *m_out, uint64_t *n_in_out, void *d_ws,
                                               size_t ws_bytes, lstc_stream s) {
    hipStream_t st = (hipStream_t)s;
    if (!n_out || !m_out || !n_in_out) return fail(LSTC_ERR_INVALID, "lstc_clean_orient_split: NULL size outputs");
    if (!d_in_off || !d_in_col || !d_col_seg || !d_in_seg)
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient_split: in-list and segment-word outputs are required");
    const uint64_t *d_sizes = nullptr;
    const uint32_t *d_bad = nullptr;
    LSTC_TRY(clean_orient_launch(d_src, d_dst, m_raw, n, compact, d_sym_off, d_sym_col, d_off, d_col, d_in_off,
                                 d_in_col, d_ws, ws_bytes, st, &d_sizes, &d_bad));
    if (m_raw) {
        const uint32_t grid = (uint32_t)hmin<uint64_t>((2 * m_raw + 255) / 256, (uint64_t)device_cus() * 32);
        hipLaunchKernelGGL(k_prep_seg_words, dim3(grid), dim3(256), 0, st, d_sizes, (const uint64_t *)d_off,
                           (const uint32_t *)d_col, (const uint32_t *)d_in_col, d_col_seg, d_in_seg);
        LSTC_LAUNCHED("k_prep_seg_words");
    }
    uint64_t h[4];
    uint32_t bad = 0;
    LSTC_HIP(hipMemcpyAsync(h, d_sizes, sizeof(h), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    if (bad) return fail(LSTC_ERR_INVALID, "lstc_clean_orient_split: vertex id >= n in input");
    *n_out = (uint32_t)h[1];
    *m_out = h[2];
    *n_in_out = h[3];
    return LSTC_OK;
}

extern "C" lstc_status lstc_clean_orient(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                                         uint32_t n, int compact, uint64_t *d_sym_off,
                                         uint32_t *d_sym_col, uint64_t *d_off, uint32_t *d_col,
                                         uint64_t *d_in_off, uint32_t *d_in_col, uint32_t *n_out,
                                         uint64_t *m_out, void *d_ws, size_t ws_bytes, lstc_stream s) {
    hipStream_t st = (hipStream_t)s;
    if (!n_out || !m_out) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL size outputs");
    const uint64_t *d_sizes = nullptr;
    const uint32_t *d_bad = nullptr;
    if ((d_in_off == nullptr) != (d_in_col == nullptr))
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient: give both in-list outputs or neither");
    LSTC_TRY(clean_orient_launch(d_src, d_dst, m_raw, n, compact, d_sym_off, d_sym_col, d_off, d_col, d_in_off,
                                 d_in_col, d_ws, ws_bytes, st, &d_sizes, &d_bad));
    uint64_t h[4];
    uint32_t bad = 0;
    LSTC_HIP(hipMemcpyAsync(h, d_sizes, sizeof(h), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    if (bad) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: vertex id >= n in input");
    *n_out = (uint32_t)h[1];
    *m_out = h[2];
    return LSTC_OK;
}
