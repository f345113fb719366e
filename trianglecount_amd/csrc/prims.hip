// prims.hip -- device primitives written for this library: fills, exclusive
// scans, and an LSD radix sort of uint64 keys (8-bit digits, wave64 match-based
// stable ranking).  They replace the CUB calls of the reference
// (cub::DeviceScan::InclusiveSum at preprocess_GPU/Halo_cleanned_preprocess/
// round_clean_subgraph.cu:561-567, cub::DeviceMergeSort/Select::Unique at
// :468-495, cp.cumsum at Med_dataset/without_part_preprocess.py:256).
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

// ------------------------------------------------------------------ fills
template <class T>
__global__ void k_fill(T *p, T v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <class T>
static lstc_status fill_impl(T *p, T v, uint64_t n, hipStream_t s, const char *name) {
    if (n == 0) return LSTC_OK;
    uint32_t grid = (uint32_t)hmin<uint64_t>((n + 255) / 256, (uint64_t)device_cus() * 16);
    hipLaunchKernelGGL(k_fill<T>, dim3(grid), dim3(256), 0, s, p, v, n);
    LSTC_LAUNCHED(name);
    return LSTC_OK;
}
lstc_status fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s) { return fill_impl(p, v, n, s, "fill_u64"); }
lstc_status fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) { return fill_impl(p, v, n, s, "fill_u32"); }
lstc_status fill_u8(uint8_t *p, uint8_t v, uint64_t n, hipStream_t s) { return fill_impl(p, v, n, s, "fill_u8"); }

// ------------------------------------------------------------------ scan
// Tile = 256 threads x 16 items.  reduce -> (recursive) scan of tile sums -> downsweep.
static constexpr int kScanThreads = 256;
static constexpr int kScanItems = 16;
static constexpr uint64_t kScanTile = kScanThreads * kScanItems;

template <class TI>
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t *s_w, uint64_t *total) {
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t incl = wave_incl_scan64(v, lane);
    if (lane == 63) s_w[wv] = incl;
    __syncthreads();
    uint64_t wp = 0, tot = 0;
    for (uint32_t i = 0; i < kScanThreads / 64; ++i) {
        uint64_t x = s_w[i];
        if (i < wv) wp += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return wp + incl - v;
}

template <class TI>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const TI *__restrict__ in, uint64_t n,
                                                              uint64_t *__restrict__ tile_sums) {
    __shared__ uint64_t s_w[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t v = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        uint64_t i = base + (uint64_t)r * kScanThreads + threadIdx.x;
        if (i < n) v += (uint64_t)in[i];
    }
    uint64_t tot;
    (void)block_excl_scan<TI>(v, s_w, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Downsweep of one tile: loads and stores are striped over the threads
// (consecutive threads, consecutive items: coalesced), and the items pass
// through LDS so that each thread scans 16 consecutive items serially.  LDS
// rows are padded by one item per 16 so the per-thread reads do not collide
// on banks.  `in` may alias `out`: every read of the tile precedes the first
// barrier and every write follows it.
template <class TI>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const TI *in, uint64_t *out, uint64_t n,
                                                            const uint64_t *__restrict__ tile_off,
                                                            uint64_t *__restrict__ d_total) {
    __shared__ uint64_t s_w[kScanThreads / 64];
    __shared__ uint64_t s_x[kScanTile + kScanTile / 16];
    const uint32_t tid = threadIdx.x;
    const uint64_t tb = (uint64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const uint32_t e = r * kScanThreads + tid;
        s_x[e + e / 16] = tb + e < n ? (uint64_t)in[tb + e] : 0;
    }
    __syncthreads();
    uint64_t x[kScanItems], v = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const uint32_t e = tid * kScanItems + r;
        x[r] = s_x[e + e / 16];
        v += x[r];
    }
    uint64_t tot;
    uint64_t run = block_excl_scan<TI>(v, s_w, &tot) + (tile_off ? tile_off[blockIdx.x] : 0);
    // block_excl_scan's barriers order every thread's reads of s_x before these writes
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const uint32_t e = tid * kScanItems + r;
        s_x[e + e / 16] = run;
        run += x[r];
    }
    if (d_total && blockIdx.x == gridDim.x - 1 && tid == kScanThreads - 1) *d_total = run;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const uint32_t e = r * kScanThreads + tid;
        if (tb + e < n) out[tb + e] = s_x[e + e / 16];
    }
}

static size_t scan_ws_rec(uint64_t n) {
    if (n <= kScanTile) return 0;
    uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    return ((tiles * sizeof(uint64_t) + 255) & ~size_t(255)) + scan_ws_rec(tiles);
}
size_t scan_ws_bytes(uint64_t n) { return scan_ws_rec(n) + 256; }

template <class TI>
static lstc_status scan_impl(const TI *in, uint64_t *out, uint64_t n, uint64_t *d_total, void *ws,
                             size_t ws_bytes, hipStream_t s) {
    if (n == 0) {
        if (d_total) return fill_u64(d_total, 0, 1, s);
        return LSTC_OK;
    }
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_down<TI>, dim3(1), dim3(kScanThreads), 0, s, in, out, n,
                           (const uint64_t *)nullptr, d_total);
        LSTC_LAUNCHED("k_scan_down");
        return LSTC_OK;
    }
    if (tiles > 0xFFFFFFFFull) return fail(LSTC_ERR_INVALID, "scan: too many items");
    size_t need = ((tiles * sizeof(uint64_t) + 255) & ~size_t(255));
    if (!ws || ws_bytes < need) return fail(LSTC_ERR_WORKSPACE, "scan: workspace too small");
    uint64_t *sums = (uint64_t *)ws;
    hipLaunchKernelGGL(k_scan_reduce<TI>, dim3((uint32_t)tiles), dim3(kScanThreads), 0, s, in, n, sums);
    LSTC_LAUNCHED("k_scan_reduce");
    LSTC_TRY(scan_impl<uint64_t>(sums, sums, tiles, nullptr, (char *)ws + need, ws_bytes - need, s));
    hipLaunchKernelGGL(k_scan_down<TI>, dim3((uint32_t)tiles), dim3(kScanThreads), 0, s, in, out, n,
                       (const uint64_t *)sums, d_total);
    LSTC_LAUNCHED("k_scan_down");
    return LSTC_OK;
}

lstc_status scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *d_total,
                               void *ws, size_t ws_bytes, hipStream_t s) {
    return scan_impl<uint64_t>(in, out, n, d_total, ws, ws_bytes, s);
}
lstc_status scan_exclusive_u32_to_u64(const uint32_t *in, uint64_t *out, uint64_t n,
                                      uint64_t *d_total, void *ws, size_t ws_bytes, hipStream_t s) {
    return scan_impl<uint32_t>(in, out, n, d_total, ws, ws_bytes, s);
}

// ------------------------------------------------------------------ radix sort
// LSD over 8-bit digits; per digit a histogram pass, one scan of the
// digit-major block counts, and a scatter pass.  A block owns kSortSub
// consecutive sub-tiles of 4 waves x 16 rounds x 64 lanes = 4096 keys; key i of
// a sub-tile sits at (wave, round, lane) = (i / 1024, (i / 64) % 16, i % 64), so
// ranking in that lexicographic order, sub-tile after sub-tile, is stable.
// Owning several sub-tiles keeps the count matrix (one count per block and
// digit, written and read at a stride) small next to the keys.  Both passes
// load a sub-tile's keys before they use any, and the next sub-tile's while
// they work on the current one.
static constexpr int kSortWaves = 4;
static constexpr int kSortRounds = 16;
static constexpr uint32_t kSortSub = 4;
static constexpr uint64_t kSortTile = kSortWaves * kSortRounds * 64;  // keys per sub-tile
static constexpr uint64_t kSortBlock = kSortTile * kSortSub;          // keys per block
static_assert(kSortBlock == kRadixBlock, "the sort's block is the one key producers count by");

// Digit counts of each block.  The keys are loaded with the non-temporal hint:
// each is read once here, and the pass measured 12% faster with it (0.40 vs
// 0.45 ms per pass on 268 M keys); the scatter's own loads and stores keep the
// default policy, where the hint measured slower.
__global__ __launch_bounds__(kSortWaves * 64) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                                uint64_t n, uint32_t shift,
                                                                uint32_t num_blocks,
                                                                uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    const uint64_t base = (uint64_t)blockIdx.x * kSortBlock;
    uint64_t k[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint64_t i = base + (uint64_t)r * (kSortWaves * 64) + threadIdx.x;
        k[r] = i < n ? __builtin_nontemporal_load(&keys[i]) : 0;
    }
    __syncthreads();
    for (uint32_t sub = 0; sub < kSortSub; ++sub) {
        const uint64_t b0 = base + sub * kSortTile;
        uint64_t kn[kSortRounds];
        if (sub + 1 < kSortSub) {
#pragma unroll
            for (int r = 0; r < kSortRounds; ++r) {
                const uint64_t i = b0 + kSortTile + (uint64_t)r * (kSortWaves * 64) + threadIdx.x;
                kn[r] = i < n ? __builtin_nontemporal_load(&keys[i]) : 0;
            }
        }
#pragma unroll
        for (int r = 0; r < kSortRounds; ++r)
            if (b0 + (uint64_t)r * (kSortWaves * 64) + threadIdx.x < n)
                atomicAdd(&cnt[(uint32_t)(k[r] >> shift) & 255u], 1u);
        if (sub + 1 < kSortSub) {
#pragma unroll
            for (int r = 0; r < kSortRounds; ++r) k[r] = kn[r];
        }
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * num_blocks + blockIdx.x] = cnt[threadIdx.x];
}

// Scatter of a block's sub-tiles, in order: keys are ranked stably per digit
// (ballot match inside a wave, running counts across rounds and waves), staged
// in LDS in digit order, then written out by consecutive threads, so each
// digit's run of the sub-tile leaves as contiguous stores; the block's running
// offset of each digit advances by the sub-tile's count.
__global__ __launch_bounds__(kSortWaves * 64) void k_radix_scatter(
    const uint64_t *__restrict__ keys, uint64_t *__restrict__ out, uint64_t n, uint32_t shift,
    uint32_t num_blocks, const uint64_t *__restrict__ goff) {
    __shared__ uint32_t wcnt[kSortWaves][256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t wsum[kSortWaves];
    __shared__ uint64_t gbase[256];
    __shared__ uint64_t stage[kSortTile];
    const uint32_t tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    uint64_t run = goff[(uint64_t)tid * num_blocks + blockIdx.x];  // digit tid's next output position
    uint64_t k[kSortRounds];
    {
        const uint64_t b = (uint64_t)blockIdx.x * kSortBlock + (uint64_t)wv * (kSortRounds * 64);
#pragma unroll
        for (int r = 0; r < kSortRounds; ++r) {
            const uint64_t i = b + (uint64_t)r * 64 + lane;
            k[r] = i < n ? keys[i] : 0;
        }
    }
    for (uint32_t sub = 0; sub < kSortSub; ++sub) {
        const uint64_t tile0 = (uint64_t)blockIdx.x * kSortBlock + sub * kSortTile;
        if (tile0 >= n) break;  // block-uniform
        const uint64_t base = tile0 + (uint64_t)wv * (kSortRounds * 64);
        uint64_t kn[kSortRounds];
        if (sub + 1 < kSortSub) {
#pragma unroll
            for (int r = 0; r < kSortRounds; ++r) {
                const uint64_t i = base + kSortTile + (uint64_t)r * 64 + lane;
                kn[r] = i < n ? keys[i] : 0;
            }
        }
        for (int w = 0; w < kSortWaves; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        uint32_t rank[kSortRounds];
#pragma unroll
        for (int r = 0; r < kSortRounds; ++r) {
            const bool valid = base + (uint64_t)r * 64 + lane < n;
            const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
            uint64_t peers = ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t bm = ballot(bit);
                peers &= bit ? bm : ~bm;
            }
            uint32_t before = 0;
            if (valid) before = wcnt[wv][d];
            rank[r] = before + mask_rank(peers);
            if (valid && mask_rank(peers) == 0) wcnt[wv][d] = before + (uint32_t)__popcll(peers);
        }
        __syncthreads();
        // thread tid owns digit tid: offsets of the waves inside the digit, the
        // digit's start inside the sub-tile, and its output base
        uint32_t tot = 0;
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t t = wcnt[w][tid];
            wcnt[w][tid] = tot;
            tot += t;
        }
        const uint32_t incl = wave_incl_scan(tot, lane);
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t w = 0; w < wv; ++w) before += wsum[w];
        const uint32_t start = before + incl - tot;
        dstart[tid] = start;
        gbase[tid] = run - start;
        run += tot;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortRounds; ++r) {
            const uint64_t i = base + (uint64_t)r * 64 + lane;
            if (i < n) {
                const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
                stage[dstart[d] + wcnt[wv][d] + rank[r]] = k[r];
            }
        }
        __syncthreads();
        const uint32_t tile_n = (uint32_t)(n - tile0 < kSortTile ? n - tile0 : kSortTile);
        for (uint32_t i = tid; i < tile_n; i += kSortWaves * 64) {
            const uint64_t key = stage[i];
            out[gbase[(uint32_t)(key >> shift) & 255u] + i] = key;
        }
        if (sub + 1 < kSortSub) {
#pragma unroll
            for (int r = 0; r < kSortRounds; ++r) k[r] = kn[r];
        }
        __syncthreads();  // stage, wcnt and gbase are reused by the next sub-tile
    }
}

static size_t radix_layout(uint64_t n, uint32_t **hist, uint64_t **goff, void **scan_ws,
                           size_t *scan_bytes, void *ws, size_t ws_bytes) {
    const uint64_t tiles = (n + kSortBlock - 1) / kSortBlock;
    Carver c(ws, ws_bytes);
    *hist = c.take<uint32_t>(tiles * 256);
    *goff = c.take<uint64_t>(tiles * 256);
    size_t sb = scan_ws_bytes(tiles * 256);
    *scan_ws = c.take<char>(sb);
    *scan_bytes = sb;
    return c.need();
}

size_t radix_sort_ws_bytes(uint64_t n) {
    uint32_t *h; uint64_t *g; void *sw; size_t sb;
    return radix_layout(n, &h, &g, &sw, &sb, nullptr, 0);
}

uint32_t *radix_first_counts(void *ws, uint64_t n) {
    uint32_t *h; uint64_t *g; void *sw; size_t sb;
    radix_layout(n, &h, &g, &sw, &sb, ws, SIZE_MAX);
    return h;
}

lstc_status radix_sort_u64(uint64_t *keys, uint64_t *alt, uint64_t n, int key_bits,
                           uint64_t **result, void *ws, size_t ws_bytes, hipStream_t s,
                           bool first_counts_ready) {
    *result = keys;
    if (n <= 1 || key_bits <= 0) return LSTC_OK;
    uint32_t *hist; uint64_t *goff; void *sws; size_t sb;
    size_t need = radix_layout(n, &hist, &goff, &sws, &sb, ws, ws_bytes);
    if (!ws || ws_bytes < need) return fail(LSTC_ERR_WORKSPACE, "radix sort: workspace too small");
    const uint64_t tiles = (n + kSortBlock - 1) / kSortBlock;
    if (tiles > 0x7FFFFFFFull) return fail(LSTC_ERR_INVALID, "radix sort: too many keys");
    uint64_t *in = keys, *out = alt;
    for (int shift = 0; shift < key_bits; shift += 8) {
        if (shift > 0 || !first_counts_ready) {
            hipLaunchKernelGGL(k_radix_hist, dim3((uint32_t)tiles), dim3(kSortWaves * 64), 0, s, in, n,
                               (uint32_t)shift, (uint32_t)tiles, hist);
            LSTC_LAUNCHED("k_radix_hist");
        }
        LSTC_TRY(scan_exclusive_u32_to_u64(hist, goff, tiles * 256, nullptr, sws, sb, s));
        hipLaunchKernelGGL(k_radix_scatter, dim3((uint32_t)tiles), dim3(kSortWaves * 64), 0, s, in,
                           out, n, (uint32_t)shift, (uint32_t)tiles, goff);
        LSTC_LAUNCHED("k_radix_scatter");
        uint64_t *t = in; in = out; out = t;
    }
    *result = in;
    return LSTC_OK;
}

}  // namespace lstc
