// prep.hip -- (a) GPU clean + orient into a uint32 CSR with uint64 offsets.
//
// Replaces the DGL cleaning of Med_dataset/without_part_preprocess.py:193-208
// (remove_self_loop, to_bidirected, isolated-node removal) and the CuPy
// orientation Find_size / cp.cumsum / Convert of :19-111,253-270.
//
// Pipeline (all stream-ordered, sizes kept on the device until the end):
//   keys  : each raw pair (s, d), s != d, emits (s<<b | d) and (d<<b | s);
//           loops and out-of-range ids emit a sentinel that sorts last;
//           the same kernel counts the sort's first digit per sort block
//   sort  : LSD radix sort over 2b bits (prims.hip)
//   then compactions of the sorted list, each a count (per tile of 4096 keys),
//   a scan of the tile counts and a scatter that recomputes the predicate and
//   writes the kept entries at their final positions (no flag or position
//   arrays):
//   unique : first of each run -> symmetric arc list, row bounds -> deg[v]
//   relabel (compact): id' = rank of v among vertices with deg > 0 (monotone)
//   symmetric CSR: offsets = scan(deg) compacted, columns relabelled
//   orient : keep s->d iff deg s < deg d or (equal and d < s); the count pass
//            also adds d+(v) per run of equal sources, so offsets = scan(d+)
//            are known before the scatter, which writes the columns and their
//            segment words (start | d+ << 32 of the target's row) from one
//            16-byte record per endpoint and counts the in-list arcs
//   in     : reverse keys of kept arcs with d+(d) <= d+(s) -> target-side
//            in-lists and their segment words
#include "lookback.h"
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

// Keys of both directions of every raw pair (pair i gives keys 2i and 2i + 1;
// loops and out-of-range ids the sentinel), one block of the radix sort per
// workgroup: the workgroup also counts its block's first digit (the low 8
// bits) into counts[d * blocks + blockIdx.x], the sort's first histogram.
// The raw pairs are read once (non-temporal hint).
static constexpr uint32_t kKeyPairs = (uint32_t)(kRadixBlock / 2);  // raw pairs per workgroup
static constexpr uint32_t kKeyBatch = 8;                             // rounds whose loads go out together

__global__ __launch_bounds__(256) void k_make_keys(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                                                   uint64_t m_raw, uint32_t n, uint32_t b, uint64_t *__restrict__ keys,
                                                   uint32_t blocks, uint32_t *__restrict__ counts,
                                                   uint32_t *__restrict__ bad) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t p0 = (uint64_t)blockIdx.x * kKeyPairs;
    bool out_of_range = false;
    for (uint32_t r0 = 0; r0 < kKeyPairs / 256; r0 += kKeyBatch) {
        uint32_t s[kKeyBatch], d[kKeyBatch];
#pragma unroll
        for (uint32_t r = 0; r < kKeyBatch; ++r) {
            const uint64_t i = p0 + (uint64_t)(r0 + r) * 256 + threadIdx.x;
            s[r] = i < m_raw ? __builtin_nontemporal_load(&src[i]) : 0;
            d[r] = i < m_raw ? __builtin_nontemporal_load(&dst[i]) : 0;
        }
#pragma unroll
        for (uint32_t r = 0; r < kKeyBatch; ++r) {
            const uint64_t i = p0 + (uint64_t)(r0 + r) * 256 + threadIdx.x;
            if (i >= m_raw) break;
            uint64_t k0 = kSentinel, k1 = kSentinel;
            if (s[r] >= n || d[r] >= n) {
                out_of_range = true;
            } else if (s[r] != d[r]) {
                k0 = ((uint64_t)s[r] << b) | d[r];
                k1 = ((uint64_t)d[r] << b) | s[r];
            }
            *(ulonglong2 *)(keys + 2 * i) = make_ulonglong2(k0, k1);
            atomicAdd(&cnt[(uint32_t)k0 & 255u], 1u);
            atomicAdd(&cnt[(uint32_t)k1 & 255u], 1u);
        }
    }
    if (out_of_range) *bad = 1;
    __syncthreads();
    counts[(uint64_t)threadIdx.x * blocks + blockIdx.x] = cnt[threadIdx.x];
}

// ------------------------------------------------------------------ compactions
// Tile = 256 threads x 16 rounds = 4096 keys; key (round r, thread t) is tile
// position r * 256 + t, so loads are coalesced and a wave's 64 lanes hold 64
// consecutive keys of a round.  A count pass gives each tile's number of kept
// keys; after a scan of those, a scatter pass writes the kept keys in order,
// round by round (one ballot per round and a running count across the
// workgroup), so no flag or position array is needed.  Loads are issued in
// batches: a thread first loads all its keys, then the gathers of kCG rounds at
// a time, so each thread has many loads in flight instead of one dependent
// chain per key.  The tile's keys also sit in LDS, with the key before and
// after the tile, for the tests that look at neighbours.  The count pass loads
// the keys with the non-temporal hint (measured a little faster: the scatter
// pass reads the 2 GB of keys again from memory either way).
//
// An Op provides: size(), keys(); fetch(k) -> Aux, the gathers its predicate
// and its output need; pred(i, k, prev, aux); emit(i, k, prev, next, kept, pos,
// aux).  prev / next are the keys at i - 1 / i + 1 when kNeedsPrev (else 0;
// unspecified at i = 0 / i = n - 1, where the Op must not use them).  kRuns:
// the count pass also adds each source's kept keys into runs()[source].
static constexpr uint32_t kCT = 256;
static constexpr uint32_t kCR = 16;
static constexpr uint32_t kCG = 8;  // rounds per batch of gathers
static constexpr uint64_t kCTile = (uint64_t)kCT * kCR;

// Adds, for each run of equal sources among the wave's lanes, the number of its
// kept keys into cnt[source] (one atomic per run; sources ascend across lanes).
__device__ __forceinline__ void add_runs(uint32_t *cnt, uint32_t src, bool kept, bool valid, uint32_t lane) {
    const uint32_t prev = (uint32_t)__shfl_up((int)src, 1);
    const uint64_t heads = ballot(valid && (lane == 0 || prev != src));
    const uint64_t km = ballot(valid && kept);
    if (valid && lane_in(heads)) {
        const uint64_t above = heads & ~((2ull << lane) - 1ull);  // heads after this lane (none at lane 63)
        const uint64_t upto = above ? ((above & (0ull - above)) - 1ull) : ~0ull;
        const uint32_t c = (uint32_t)__popcll(km & upto & ~((1ull << lane) - 1ull));
        if (c) atomicAdd(cnt + src, c);
    }
}

template <class Op>
__global__ __launch_bounds__(kCT) void k_compact_count(Op op, uint32_t *__restrict__ tile_cnt) {
    __shared__ uint64_t s_k[Op::kNeedsPrev ? kCTile + 1 : 1];  // s_k[p + 1] = key at tile position p
    __shared__ uint32_t s_w[kCT / 64];
    const uint64_t n = op.size();
    const uint64_t tile0 = (uint64_t)blockIdx.x * kCTile;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t c = 0;
    if (tile0 < n) {
        const uint64_t *keys = op.keys();
        uint64_t k[kCR];
#pragma unroll
        for (uint32_t r = 0; r < kCR; ++r) {
            const uint64_t i = tile0 + (uint64_t)r * kCT + threadIdx.x;
            k[r] = i < n ? __builtin_nontemporal_load(&keys[i]) : 0;
        }
        if (Op::kNeedsPrev) {
#pragma unroll
            for (uint32_t r = 0; r < kCR; ++r) s_k[1 + r * kCT + threadIdx.x] = k[r];
            if (threadIdx.x == 0) s_k[0] = tile0 ? keys[tile0 - 1] : 0;
            __syncthreads();
        }
#pragma unroll
        for (uint32_t g = 0; g < kCR; g += kCG) {
            typename Op::Aux a[kCG];
#pragma unroll
            for (uint32_t r = 0; r < kCG; ++r) a[r] = op.fetch(k[g + r]);
#pragma unroll
            for (uint32_t r = 0; r < kCG; ++r) {
                const uint32_t p = (g + r) * kCT + threadIdx.x;
                const uint64_t i = tile0 + p;
                const bool kp = i < n && op.pred(i, k[g + r], Op::kNeedsPrev ? s_k[p] : 0, a[r]);
                c += kp ? 1u : 0u;
                if (Op::kRuns) add_runs(op.runs(), op.src(k[g + r]), kp, i < n, lane);
            }
        }
    }
    c = wave_sum32(c);
    if (lane == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

template <class Op>
__global__ __launch_bounds__(kCT) void k_compact_scatter(Op op, const uint64_t *__restrict__ tile_off) {
    __shared__ uint64_t s_k[Op::kNeedsPrev ? kCTile + 2 : 1];  // s_k[p + 1] = key at position p; s_k[0], s_k[kCTile + 1]
    __shared__ uint32_t s_w[2][kCT / 64];    // kept keys per wave of a round (two rounds in flight)
    const uint64_t n = op.size();
    const uint64_t tile0 = (uint64_t)blockIdx.x * kCTile;
    if (tile0 >= n) return;  // the whole block leaves together
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t *keys = op.keys();
    uint64_t k[kCR];
#pragma unroll
    for (uint32_t r = 0; r < kCR; ++r) {
        const uint64_t i = tile0 + (uint64_t)r * kCT + threadIdx.x;
        k[r] = i < n ? keys[i] : 0;
    }
    if (Op::kNeedsPrev) {
#pragma unroll
        for (uint32_t r = 0; r < kCR; ++r) s_k[1 + r * kCT + threadIdx.x] = k[r];
        if (threadIdx.x == 0) s_k[0] = tile0 ? keys[tile0 - 1] : 0;
        if (threadIdx.x == 1) s_k[kCTile + 1] = tile0 + kCTile < n ? keys[tile0 + kCTile] : 0;
        __syncthreads();
    }
    uint64_t pos = tile_off[blockIdx.x];  // output position of the round's first kept key
#pragma unroll
    for (uint32_t g = 0; g < kCR; g += kCG) {
        typename Op::Aux a[kCG];
#pragma unroll
        for (uint32_t r = 0; r < kCG; ++r) a[r] = op.fetch(k[g + r]);
#pragma unroll
        for (uint32_t r = 0; r < kCG; ++r) {
            const uint32_t p = (g + r) * kCT + threadIdx.x;
            const uint64_t i = tile0 + p;
            const bool valid = i < n;
            const uint64_t prev = Op::kNeedsPrev ? s_k[p] : 0, next = Op::kNeedsPrev ? s_k[p + 2] : 0;
            const bool kp = valid && op.pred(i, k[g + r], prev, a[r]);
            const uint64_t km = ballot(kp);
            if (lane == 0) s_w[r & 1][wv] = (uint32_t)__popcll(km);
            __syncthreads();
            uint32_t before = 0, total = 0;
#pragma unroll
            for (uint32_t x = 0; x < kCT / 64; ++x) {
                const uint32_t v = s_w[r & 1][x];
                before += x < wv ? v : 0u;
                total += v;
            }
            if (valid) op.emit(i, k[g + r], prev, next, kp, pos + before + mask_rank(km), a[r]);
            pos += total;
        }
    }
}

// Unique keys of the sorted list (sentinels dropped), with each row's bounds
// [first, last) in the unique list.
struct UniqueOp {
    const uint64_t *sorted;
    uint64_t n;
    uint32_t b;
    uint64_t *uk, *first, *last;
    struct Aux {};
    static constexpr bool kNeedsPrev = true, kRuns = false;
    __device__ uint64_t size() const { return n; }
    __device__ const uint64_t *keys() const { return sorted; }
    __device__ Aux fetch(uint64_t) const { return Aux{}; }
    __device__ bool pred(uint64_t i, uint64_t k, uint64_t prev, Aux) const {
        return k != kSentinel && (i == 0 || prev != k);
    }
    __device__ uint32_t src(uint64_t k) const { return (uint32_t)(k >> b); }
    __device__ uint32_t *runs() const { return nullptr; }
    __device__ void emit(uint64_t i, uint64_t k, uint64_t prev, uint64_t next, bool kept, uint64_t pos, Aux) const {
        if (k == kSentinel) return;
        if (kept) uk[pos] = k;
        const uint64_t s = k >> b;
        if (i == 0 || (prev >> b) != s) first[s] = pos;  // a row's first key is always kept
        if (i + 1 == n || (next >> b) != s) last[s] = pos + (kept ? 1u : 0u);
    }
};

// The orientation rule: keep s -> d iff deg s < deg d, or equal and d < s.
__device__ __forceinline__ bool out_arc(uint32_t ds, uint32_t dd, uint32_t s, uint32_t d) {
    return ds < dd || (ds == dd && d < s);
}

// Count pass of the oriented arcs over the unique symmetric keys (s << b | d,
// old ids): kept arcs per tile, and d+ per old source id.
struct OrientCountOp {
    const uint64_t *uk;
    const uint64_t *d_m2;
    uint32_t b;
    uint64_t mask;
    const uint32_t *deg;  // symmetric degree per old id
    uint32_t *dout;
    struct Aux { uint32_t ds, dd; };
    static constexpr bool kNeedsPrev = false, kRuns = true;
    __device__ uint64_t size() const { return *d_m2; }
    __device__ const uint64_t *keys() const { return uk; }
    __device__ Aux fetch(uint64_t k) const { return Aux{deg[(uint32_t)(k >> b)], deg[(uint32_t)(k & mask)]}; }
    __device__ bool pred(uint64_t, uint64_t k, uint64_t, Aux a) const {
        return out_arc(a.ds, a.dd, (uint32_t)(k >> b), (uint32_t)(k & mask));
    }
    __device__ uint32_t src(uint64_t k) const { return (uint32_t)(k >> b); }
    __device__ uint32_t *runs() const { return dout; }
    __device__ void emit(uint64_t, uint64_t, uint64_t, uint64_t, bool, uint64_t, Aux) const {}
};

// Per old vertex id: its new id, d+, the start of its oriented row (low 32
// bits) and its symmetric degree -- everything a kept arc's endpoints add to
// the predicates and to the outputs, in one 16-byte gather.
struct __align__(16) VRec {
    uint32_t nid, dplus, off_lo, deg;
};

__device__ __forceinline__ uint64_t seg_word(const VRec &v) { return (uint64_t)v.off_lo | ((uint64_t)v.dplus << 32); }

// An arc the split count handles from the target side: in row s, the source d
// of an oriented arc d -> s with d+(d) <= d+(s).
__device__ __forceinline__ bool in_arc(const VRec &vs, const VRec &vd, uint32_t s, uint32_t d) {
    return !out_arc(vs.deg, vd.deg, s, d) && vd.dplus <= vs.dplus;
}

// Scatter pass of the oriented arcs and of the in-list arcs together, over the
// unique symmetric keys (old ids), one tile of 4096 keys per workgroup:
//   oriented arcs: the relabelled target and, if asked, its segment word
//     (start | d+ << 32 of the target's row), at positions from the count pass's
//     tile offsets;
//   in-list arcs (in row s, the source d of an oriented arc d -> s with
//     d+(d) <= d+(s)): the source and its segment word, at positions from a
//     chained scan across tiles (decoupled look-back), and their number per
//     source for the in-list offsets.
// Tiles are numbered by a ticket taken when a workgroup starts, so every earlier
// tile belongs to a workgroup that is running or done, and that workgroup
// publishes its count before it waits on anything: the look-back ends.  Its
// wait is also bounded; past the bound bit 1 of *flag is set and the in-lists
// are not valid (the call then fails).
struct OrientInPass {
    const uint64_t *uk;
    const uint64_t *d_m2;
    uint32_t b;
    uint64_t mask;
    const VRec *rec;
    uint32_t *col;
    uint64_t *col_seg;
    uint32_t *in_col;   // NULL: no in-lists
    uint64_t *in_seg;
    uint32_t *din;
    uint64_t *status;   // per tile: flag << 62 | count, zero = not published
    uint32_t *ticket;
    uint64_t *in_total;
    uint32_t *flag;
    uint32_t *scol;           // streaming copy (NULL: none): row s from rec[s].off_lo, 128-byte aligned
    const uint64_t *doscan;   // row starts in col (old ids), for an arc's index in its row
};

__global__ __launch_bounds__(kCT) void k_orient_in_scatter(OrientInPass op, const uint64_t *__restrict__ tile_off) {
    __shared__ uint32_t s_w[2][kCT / 64];        // oriented arcs per wave of a round (two rounds in flight)
    __shared__ uint32_t s_in[kCR * (kCT / 64)];  // in-list arcs per (round, wave), then their exclusive prefix
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(op.ticket, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    const uint64_t n = *op.d_m2;
    const uint64_t tile0 = tile * kCTile;
    if (tile0 >= n) return;  // no keys; every later tile has none either
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool with_in = op.in_col != nullptr;
    uint64_t k[kCR];
#pragma unroll
    for (uint32_t r = 0; r < kCR; ++r) {
        const uint64_t i = tile0 + (uint64_t)r * kCT + threadIdx.x;
        k[r] = i < n ? op.uk[i] : 0;
    }
    uint64_t pos = tile_off[tile];         // output position of the round's first oriented arc
    uint64_t inm[kCR];                     // in-list arcs of each round (wave mask)
    uint32_t iv[kCR], ilo[kCR], idp[kCR];  // their source's new id, row start (low bits) and d+
#pragma unroll
    for (uint32_t g = 0; g < kCR; g += kCG) {
        VRec vs[kCG], vd[kCG];
#pragma unroll
        for (uint32_t r = 0; r < kCG; ++r) {
            vs[r] = op.rec[(uint32_t)(k[g + r] >> op.b)];
            vd[r] = op.rec[(uint32_t)(k[g + r] & op.mask)];
        }
#pragma unroll
        for (uint32_t r = 0; r < kCG; ++r) {
            const uint64_t i = tile0 + (uint64_t)(g + r) * kCT + threadIdx.x;
            const bool valid = i < n;
            const uint32_t s = (uint32_t)(k[g + r] >> op.b), d = (uint32_t)(k[g + r] & op.mask);
            const bool ko = valid && out_arc(vs[r].deg, vd[r].deg, s, d);
            const bool ki = valid && with_in && in_arc(vs[r], vd[r], s, d);
            const uint64_t km = ballot(ko);
            if (lane == 0) s_w[r & 1][wv] = (uint32_t)__popcll(km);
            inm[g + r] = ballot(ki);
            if (lane == 0) s_in[(g + r) * (kCT / 64) + wv] = (uint32_t)__popcll(inm[g + r]);
            iv[g + r] = vd[r].nid;
            ilo[g + r] = vd[r].off_lo;
            idp[g + r] = vd[r].dplus;
            if (with_in) add_runs(op.din, s, ki, valid, lane);
            __syncthreads();
            uint32_t before = 0, total = 0;
#pragma unroll
            for (uint32_t x = 0; x < kCT / 64; ++x) {
                const uint32_t v = s_w[r & 1][x];
                before += x < wv ? v : 0u;
                total += v;
            }
            if (ko) {
                const uint64_t q = pos + before + mask_rank(km);
                op.col[q] = vd[r].nid;
                if (op.col_seg) op.col_seg[q] = seg_word(vd[r]);
                if (op.scol) op.scol[vs[r].off_lo + (q - op.doscan[s])] = vd[r].nid;
            }
            pos += total;
        }
    }
    if (!with_in) return;
    __syncthreads();
    if (wv == 0) {
        // the tile's in-list arcs: exclusive prefix over (round, wave), then the
        // tile's place among all tiles
        const uint32_t v = s_in[lane];
        const uint32_t incl = wave_incl_scan(v, lane);
        s_in[lane] = incl - v;
        const uint64_t c = read_lane(incl, 63);
        const uint64_t excl = tile_lookback(op.status, tile, c, op.flag, lane);
        if (lane == 0) {
            s_excl = excl;
            if (tile0 + kCTile >= n) *op.in_total = excl + c;  // the last tile holds the total
        }
    }
    __syncthreads();
    const uint64_t base = s_excl;
#pragma unroll
    for (uint32_t r = 0; r < kCR; ++r) {
        if (!lane_in(inm[r])) continue;
        const uint64_t q = base + s_in[r * (kCT / 64) + wv] + mask_rank(inm[r]);
        op.in_col[q] = iv[r];
        if (op.in_seg) op.in_seg[q] = (uint64_t)ilo[r] | ((uint64_t)idp[r] << 32);
    }
}

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

// Row lengths from the unique pass's row bounds (the symmetric degree).
__global__ void k_row_lengths(const uint64_t *__restrict__ first, const uint64_t *__restrict__ last, uint32_t n,
                              uint32_t *__restrict__ len) {
    GRID_STRIDE(v, n) len[v] = (uint32_t)(last[v] - first[v]);
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

// Row lengths of the streaming copy: d+ rounded up to 32 entries (128 bytes).
__global__ void k_pad_rows(const uint32_t *__restrict__ dout, uint32_t n, uint32_t *__restrict__ plen) {
    GRID_STRIDE(v, n) plen[v] = (dout[v] + 31u) & ~31u;
}

// Per-vertex records for the oriented and in-list scatters (old ids; doscan is
// the exclusive scan of d+ over old ids, which is the oriented row start of the
// vertex's new id, since the relabelling keeps the order).
__global__ void k_vrec(const uint32_t *__restrict__ nid, const uint32_t *__restrict__ dout,
                       const uint64_t *__restrict__ doscan, const uint32_t *__restrict__ deg, uint32_t n,
                       VRec *__restrict__ rec) {
    GRID_STRIDE(v, n) rec[v] = VRec{nid[v], dout[v], (uint32_t)doscan[v], deg[v]};
}

struct PrepLayout {
    uint64_t *keys, *alt;
    uint32_t *tile_cnt;
    uint64_t *tile_off;
    uint64_t *tstat;  // look-back words of the in-list scan, one per tile
    uint32_t *ticket;
    uint32_t *deg, *keepv, *nid, *dout, *din;
    VRec *rec;
    uint64_t *degscan, *nidscan, *doscan, *first, *last, *discan;
    uint32_t *plen;   // streaming copy: padded row lengths, and their scan
    uint64_t *pscan;
    uint64_t *scal;  // [0]=m2 [1]=n_out [2]=m [3]=in-list entries [4]=streaming copy entries
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
    L.tstat = c.take<uint64_t>(tiles);
    L.ticket = c.take<uint32_t>(4);
    L.deg = c.take<uint32_t>((size_t)n + 1);
    L.keepv = c.take<uint32_t>((size_t)n + 1);
    L.nid = c.take<uint32_t>((size_t)n + 1);
    L.dout = c.take<uint32_t>((size_t)n + 1);
    L.din = c.take<uint32_t>((size_t)n + 1);
    L.rec = c.take<VRec>((size_t)n + 1);
    L.discan = c.take<uint64_t>((size_t)n + 1);
    L.degscan = c.take<uint64_t>((size_t)n + 1);
    L.nidscan = c.take<uint64_t>((size_t)n + 1);
    L.doscan = c.take<uint64_t>((size_t)n + 1);
    L.first = c.take<uint64_t>((size_t)n + 1);
    L.last = c.take<uint64_t>((size_t)n + 1);
    L.plen = c.take<uint32_t>((size_t)n + 1);
    L.pscan = c.take<uint64_t>((size_t)n + 1);
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
// at device words {m2, n_out, m, in-list entries} and *d_bad at the out-of-range
// flag; both are valid once the stream reaches this point.
lstc_status lstc::clean_orient_launch(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw, uint32_t n,
                                      int compact, uint64_t *d_sym_off, uint32_t *d_sym_col, uint64_t *d_off,
                                      uint32_t *d_col, uint64_t *d_in_off, uint32_t *d_in_col, uint64_t *d_col_seg,
                                      uint64_t *d_in_seg, void *d_ws, size_t ws_bytes, hipStream_t st,
                                      const uint64_t **d_sizes, const uint32_t **d_bad, uint32_t *d_scol) {
    if (n == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: n must be < 2^32-1");
    if (!d_off) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL output");
    if ((d_sym_off == nullptr) != (d_sym_col == nullptr))
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient: give both symmetric outputs or neither");
    if (m_raw && (!d_src || !d_dst || !d_col)) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL edge arrays");
    if (d_in_seg && !d_in_col) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: in-list segment words need in-lists");
    if (d_scol && (!d_col_seg || (uint64_t)m_raw + 31ull * n >= (1ull << 32)))
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient: a streaming copy needs segment words and m_raw + 31 n < 2^32");
    PrepLayout L = prep_layout(d_ws, ws_bytes, n, m_raw);
    if (!d_ws || ws_bytes < L.bytes)
        return fail(LSTC_ERR_WORKSPACE, "lstc_clean_orient: workspace %zu < %zu bytes", ws_bytes, L.bytes);
    const uint32_t b = (uint32_t)bits_for(n ? n : 1);
    const uint64_t mask = (b >= 64) ? ~0ull : ((1ull << b) - 1);
    const uint64_t nk = 2 * m_raw;
    const uint64_t tiles = (nk + kCTile - 1) / kCTile;
    const uint32_t grid_k = (uint32_t)hmin<uint64_t>((nk + 255) / 256 + 1, (uint64_t)device_cus() * 32);
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n + 255) / 256 + 1, (uint64_t)device_cus() * 32);

    LSTC_TRY(fill_u64(L.scal, 0, 8, st));
    LSTC_TRY(fill_u32(L.bad, 0, 4, st));
    LSTC_TRY(fill_u32(L.deg, 0, (uint64_t)n + 1, st));
    LSTC_TRY(fill_u32(L.dout, 0, (uint64_t)n + 1, st));
    LSTC_TRY(fill_u32(L.din, 0, (uint64_t)n + 1, st));
    if (m_raw) {
        const uint64_t kblocks = radix_blocks(nk);
        if (kblocks > 0x7FFFFFFFull) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: too many edges");
        hipLaunchKernelGGL(k_make_keys, dim3((uint32_t)kblocks), dim3(256), 0, st, d_src, d_dst, m_raw, n, b, L.keys,
                           (uint32_t)kblocks, radix_first_counts(L.sort_ws, nk), L.bad);
        LSTC_LAUNCHED("k_make_keys");
        uint64_t *sorted = nullptr;
        LSTC_TRY(radix_sort_u64(L.keys, L.alt, nk, (int)(2 * b > 64 ? 64 : 2 * b), &sorted, L.sort_ws, L.sort_bytes, st,
                                true));
        uint64_t *uk = (sorted == L.keys) ? L.alt : L.keys;
        // unique arcs and the symmetric row bounds (m2 -> scal[0])
        LSTC_TRY(fill_u64(L.first, 0, (uint64_t)n + 1, st));
        LSTC_TRY(fill_u64(L.last, 0, (uint64_t)n + 1, st));
        LSTC_TRY(compact_pass(UniqueOp{sorted, nk, b, uk, L.first, L.last}, nk, L.tile_cnt, L.tile_off, L.scal + 0,
                              L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_row_lengths, dim3(grid_n), dim3(256), 0, st, L.first, L.last, n, L.deg);
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
        // oriented arcs: count pass (tile counts, d+ per vertex; m -> scal[2]), offsets, per-vertex records
        hipLaunchKernelGGL(k_compact_count<OrientCountOp>, dim3((uint32_t)tiles), dim3(kCT), 0, st,
                           OrientCountOp{uk, L.scal + 0, b, mask, L.deg, L.dout}, L.tile_cnt);
        LSTC_LAUNCHED("k_compact_count<orient>");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.tile_cnt, L.tile_off, tiles, L.scal + 2, L.scan_ws, L.scan_bytes, st));
        LSTC_TRY(scan_exclusive_u32_to_u64(L.dout, L.doscan, n, nullptr, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                           L.scal + 1, L.scal + 2, d_off);
        LSTC_LAUNCHED("k_orient_offsets");
        if (d_scol) {  // the streaming copy's row starts (entries -> scal[4]); the records carry them
            hipLaunchKernelGGL(k_pad_rows, dim3(grid_n), dim3(256), 0, st, L.dout, n, L.plen);
            LSTC_LAUNCHED("k_pad_rows");
            LSTC_TRY(scan_exclusive_u32_to_u64(L.plen, L.pscan, n, L.scal + 4, L.scan_ws, L.scan_bytes, st));
        }
        hipLaunchKernelGGL(k_vrec, dim3(grid_n), dim3(256), 0, st, L.nid, L.dout, d_scol ? L.pscan : L.doscan, L.deg,
                           n, L.rec);
        LSTC_LAUNCHED("k_vrec");
        // scatter pass: columns (and segment words); in-lists (entries -> scal[3]) in the same pass
        LSTC_TRY(fill_u64(L.tstat, 0, tiles, st));
        LSTC_TRY(fill_u32(L.ticket, 0, 1, st));
        hipLaunchKernelGGL(k_orient_in_scatter, dim3((uint32_t)tiles), dim3(kCT), 0, st,
                           OrientInPass{uk, L.scal + 0, b, mask, L.rec, d_col, d_col_seg, d_in_col, d_in_seg, L.din,
                                        L.tstat, L.ticket, L.scal + 3, L.bad, d_scol, L.doscan},
                           (const uint64_t *)L.tile_off);
        LSTC_LAUNCHED("k_orient_in_scatter");
        if (d_in_off) {
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

// The split preparation (and, with d_scol, its streaming copy): one
// synchronisation at the end, sizes to the host.
static lstc_status clean_orient_split_impl(const char *who, const uint32_t *d_src, const uint32_t *d_dst,
                                           uint64_t m_raw, uint32_t n, int compact, uint64_t *d_sym_off,
                                           uint32_t *d_sym_col, uint64_t *d_off, uint32_t *d_col, uint64_t *d_in_off,
                                           uint32_t *d_in_col, uint64_t *d_col_seg, uint64_t *d_in_seg,
                                           uint32_t *d_scol, uint32_t *n_out, uint64_t *m_out, uint64_t *n_in_out,
                                           uint64_t *scol_out, void *d_ws, size_t ws_bytes, hipStream_t st) {
    if (!n_out || !m_out || !n_in_out) return fail(LSTC_ERR_INVALID, "%s: NULL size outputs", who);
    if (!d_in_off || !d_in_col || !d_col_seg || !d_in_seg)
        return fail(LSTC_ERR_INVALID, "%s: in-list and segment-word outputs are required", who);
    const uint64_t *d_sizes = nullptr;
    const uint32_t *d_bad = nullptr;
    LSTC_TRY(clean_orient_launch(d_src, d_dst, m_raw, n, compact, d_sym_off, d_sym_col, d_off, d_col, d_in_off,
                                 d_in_col, d_col_seg, d_in_seg, d_ws, ws_bytes, st, &d_sizes, &d_bad, d_scol));
    uint64_t h[5];
    uint32_t bad = 0;
    LSTC_HIP(hipMemcpyAsync(h, d_sizes, sizeof(h), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    if (bad & 2u) return fail(LSTC_ERR_DEVICE, "%s: in-list look-back exceeded its bound", who);
    if (bad) return fail(LSTC_ERR_INVALID, "%s: vertex id >= n in input", who);
    *n_out = (uint32_t)h[1];
    *m_out = h[2];
    *n_in_out = h[3];
    if (scol_out) *scol_out = h[4];
    return LSTC_OK;
}

extern "C" lstc_status lstc_clean_orient_split(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                                               uint32_t n, int compact, uint64_t *d_sym_off, uint32_t *d_sym_col,
                                               uint64_t *d_off, uint32_t *d_col, uint64_t *d_in_off,
                                               uint32_t *d_in_col, uint64_t *d_col_seg, uint64_t *d_in_seg,
                                               uint32_t *n_out, uint64_t *m_out, uint64_t *n_in_out, void *d_ws,
                                               size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_clean_orient_split");
    return clean_orient_split_impl("lstc_clean_orient_split", d_src, d_dst, m_raw, n, compact, d_sym_off, d_sym_col,
                                   d_off, d_col, d_in_off, d_in_col, d_col_seg, d_in_seg, nullptr, n_out, m_out,
                                   n_in_out, nullptr, d_ws, ws_bytes, (hipStream_t)s);
}

extern "C" uint64_t lstc_stream_copy_capacity(uint32_t n, uint64_t m_raw) { return m_raw + 31ull * n; }

extern "C" lstc_status lstc_clean_orient_stream(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                                                uint32_t n, int compact, uint64_t *d_sym_off, uint32_t *d_sym_col,
                                                uint64_t *d_off, uint32_t *d_col, uint64_t *d_in_off,
                                                uint32_t *d_in_col, uint64_t *d_col_seg, uint64_t *d_in_seg,
                                                uint32_t *d_scol, uint32_t *n_out, uint64_t *m_out,
                                                uint64_t *n_in_out, uint64_t *scol_out, void *d_ws, size_t ws_bytes,
                                                lstc_stream s) {
    TraceRange range_("lstc_clean_orient_stream");
    if (!d_scol || !scol_out) return fail(LSTC_ERR_INVALID, "lstc_clean_orient_stream: NULL streaming copy output");
    return clean_orient_split_impl("lstc_clean_orient_stream", d_src, d_dst, m_raw, n, compact, d_sym_off, d_sym_col,
                                   d_off, d_col, d_in_off, d_in_col, d_col_seg, d_in_seg, d_scol, n_out, m_out,
                                   n_in_out, scol_out, d_ws, ws_bytes, (hipStream_t)s);
}

extern "C" lstc_status lstc_clean_orient(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                                         uint32_t n, int compact, uint64_t *d_sym_off,
                                         uint32_t *d_sym_col, uint64_t *d_off, uint32_t *d_col,
                                         uint64_t *d_in_off, uint32_t *d_in_col, uint32_t *n_out,
                                         uint64_t *m_out, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_clean_orient");
    hipStream_t st = (hipStream_t)s;
    if (!n_out || !m_out) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL size outputs");
    const uint64_t *d_sizes = nullptr;
    const uint32_t *d_bad = nullptr;
    if ((d_in_off == nullptr) != (d_in_col == nullptr))
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient: give both in-list outputs or neither");
    LSTC_TRY(clean_orient_launch(d_src, d_dst, m_raw, n, compact, d_sym_off, d_sym_col, d_off, d_col, d_in_off,
                                 d_in_col, nullptr, nullptr, d_ws, ws_bytes, st, &d_sizes, &d_bad));
    uint64_t h[4];
    uint32_t bad = 0;
    LSTC_HIP(hipMemcpyAsync(h, d_sizes, sizeof(h), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    if (bad & 2u) return fail(LSTC_ERR_DEVICE, "lstc_clean_orient: in-list look-back exceeded its bound");
    if (bad) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: vertex id >= n in input");
    *n_out = (uint32_t)h[1];
    *m_out = h[2];
    return LSTC_OK;
}
