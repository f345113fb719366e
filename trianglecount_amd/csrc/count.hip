// count.hip -- exact triangle counting over an oriented CSR (gfx950, wave64).
//
// Semantics follow Find_Triangle + cub::DeviceReduce::Sum of
// src/WithoutPartition/ShareKernelTCV62_update_v1.cu:43-191,264,271-277: the
// count is sum over owned rows u of sum_{w in N+(u)} |N+(u) n N+(w)|.
//
// Design (one wave per work item = one owned row u):
//   1. N+(u) is inserted into a wave-private open-addressing hash table in LDS
//      (capacity = table_scale * keys; rows too long for the table are cut into
//      tiles that each re-stream the neighbour lists, like the reference's
//      128-entry shared tiles at :100-173).
//   2. The neighbour lists N+(w), w in N+(u), are walked as one flat sequence,
//      64 entries per step: a window of 64 segments is loaded (one per lane),
//      prefix-summed, empty segments are compacted out, and each lane finds its
//      segment by a 64-bit boundary mask (ballot + popcount).  A step inside a
//      single long segment skips the mask (scalar fast path).
//   3. Each entry x is probed in the LDS table; hits are counted with
//      ballot + popcount into a scalar accumulator, and each workgroup issues
//      a single uint64 atomic at the end.
// Rows are binned by d+(u) so the LDS table of each launch fits its bin.
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint32_t kEmpty = 0xFFFFFFFFu;
static constexpr int kBins = 4;

__device__ __forceinline__ uint32_t hslot(uint32_t x, uint32_t shift) {
    return (x * 0x9E3779B1u) >> shift;
}

// ------------------------------------------------------------------ binning
// bin b holds owned rows with bin_lo[b] <= d+(u) < bin_lo[b+1].
struct BinBounds {
    uint32_t lo[kBins + 1];
};

__global__ void k_count_init(uint32_t *bin_counts, unsigned long long *out) {
    if (threadIdx.x < kBins) bin_counts[threadIdx.x] = 0;
    if (threadIdx.x == 0) *out = 0;
}

__global__ __launch_bounds__(256) void k_bin_rows(const uint64_t *__restrict__ off, uint32_t n_owned,
                                                  BinBounds bb, uint32_t *__restrict__ bins,
                                                  uint32_t cap, uint32_t *__restrict__ bin_counts) {
    const uint32_t lane = lane_id();
    for (uint64_t u0 = (uint64_t)blockIdx.x * blockDim.x; u0 < n_owned;
         u0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t u = u0 + threadIdx.x;
        int b = -1;
        if (u < n_owned) {
            const uint64_t k = off[u + 1] - off[u];
            for (int i = 0; i < kBins; ++i)
                if (k >= bb.lo[i] && k < bb.lo[i + 1]) b = i;
        }
#pragma unroll
        for (int i = 0; i < kBins; ++i) {
            const uint64_t m = ballot(b == i);
            if (m == 0) continue;
            uint32_t base = 0;
            if (mask_rank(m) == 0 && b == i) base = atomicAdd(bin_counts + i, (uint32_t)__popcll(m));
            // broadcast from the first active lane of this bin
            const uint32_t first = (uint32_t)__builtin_ctzll(m);
            base = read_lane(base, first);
            if (b == i) bins[(size_t)i * cap + base + mask_rank(m)] = (uint32_t)u;
        }
        (void)lane;
    }
}

// ------------------------------------------------------------------ count kernel
template <int TLOG2, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_count_rows(
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ col,
    const uint32_t *__restrict__ items, const uint32_t *__restrict__ n_items_ptr,
    unsigned long long *__restrict__ out, unsigned long long *__restrict__ per_row,
    uint32_t tile_keys, uint32_t n_rows) {
    constexpr uint32_t CAP = 1u << TLOG2;
    constexpr uint32_t SHIFT = 32 - TLOG2;
    __shared__ uint32_t s_tab[WAVES][CAP];
    __shared__ uint64_t s_base[WAVES][64];
    __shared__ uint32_t s_end[WAVES][64];
    __shared__ uint32_t s_flag[WAVES][64];
    __shared__ unsigned long long s_sum[WAVES];

    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t *tab = s_tab[wv];
    uint32_t *flag = s_flag[wv];
    const uint32_t n_items = *n_items_ptr;
    unsigned long long acc = 0;
    flag[lane] = 0;

    for (uint32_t it = blockIdx.x * WAVES + wv; it < n_items; it += gridDim.x * WAVES) {
        const uint32_t u = uniform(items[it]);
        const uint64_t rb = off[u];
        const uint32_t K = uniform((uint32_t)(off[u + 1] - rb));
        unsigned long long rowcnt = 0;
        for (uint32_t t0 = 0; t0 < K; t0 += tile_keys) {
            const uint32_t tk = min(K - t0, tile_keys);
            for (uint32_t i = lane; i < CAP; i += 64) tab[i] = kEmpty;
            for (uint32_t j = lane; j < tk; j += 64) {
                const uint32_t key = col[rb + t0 + j];
                uint32_t h = hslot(key, SHIFT);
                while (atomicCAS(&tab[h], kEmpty, key) != kEmpty) h = (h + 1) & (CAP - 1);
            }
            for (uint32_t j0 = 0; j0 < K; j0 += 64) {
                const uint32_t j = j0 + lane;
                uint64_t s = 0;
                uint32_t len = 0;
                if (j < K) {
                    const uint32_t w = col[rb + j];
                    if (w < n_rows) {  // ids outside the CSR have no row: nothing to read
                        s = off[w];
                        len = (uint32_t)(off[w + 1] - s);
                    }
                }
                const uint32_t incl = wave_incl_scan(len, lane);
                const uint32_t total = uniform(gather(incl, 63));
                const uint64_t nzm = ballot(len != 0);
                const uint32_t nseg = (uint32_t)__popcll(nzm);
                if (len != 0) {
                    const uint32_t slot = mask_rank(nzm);
                    s_base[wv][slot] = s - (uint64_t)(incl - len);
                    s_end[wv][slot] = incl;
                }
                const uint64_t mybase = s_base[wv][lane];
                const uint32_t myend = lane < nseg ? s_end[wv][lane] : kEmpty;
                uint32_t cur = 0;
                for (uint32_t B = 0; B < total; B += 64) {
                    const uint32_t cend = read_lane(myend, cur);
                    const uint32_t t = B + lane;
                    uint64_t base;
                    if (cend >= B + 64) {
                        base = read_lane64(mybase, cur);
                    } else {
                        const bool mk = lane >= cur && myend < B + 64;
                        if (mk) flag[myend - B] = 1;
                        const uint64_t M = ballot(flag[lane] != 0);
                        if (mk) flag[myend - B] = 0;
                        const uint32_t seg = cur + (uint32_t)__popcll(M & ((2ull << lane) - 1ull));
                        base = gather64(mybase, seg);
                    }
                    bool hit = false;
                    if (t < total) {
                        const uint32_t x = col[base + t];
                        uint32_t h = hslot(x, SHIFT);
                        while (true) {
                            const uint32_t v = tab[h];
                            if (v == x) { hit = true; break; }
                            if (v == kEmpty) break;
                            h = (h + 1) & (CAP - 1);
                        }
                    }
                    rowcnt += (unsigned long long)__popcll(ballot(hit));
                    cur += (uint32_t)__popcll(ballot(lane >= cur && myend <= B + 64));
                }
            }
        }
        acc += rowcnt;
        if (per_row != nullptr && lane == 0) per_row[u] = rowcnt;
    }
    if (lane == 0) s_sum[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < WAVES; ++i) t += s_sum[i];
        if (t) atomicAdd(out, t);
    }
}

// ------------------------------------------------------------------ host side
struct CountLayout {
    uint32_t *bin_counts;
    uint32_t *bins;
    size_t bytes;
};

static CountLayout count_layout(void *ws, size_t ws_bytes, uint32_t n_owned) {
    Carver c(ws, ws_bytes);
    CountLayout L;
    L.bin_counts = c.take<uint32_t>(64);
    L.bins = c.take<uint32_t>((size_t)kBins * (n_owned ? n_owned : 1));
    L.bytes = c.need();
    return L;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_count_workspace(uint32_t n_rows, uint32_t n_owned, uint64_t m,
                                            const lstc_tune *tune, size_t *bytes) {
    (void)n_rows; (void)m; (void)tune;
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_count_workspace: bytes is NULL");
    *bytes = count_layout(nullptr, 0, n_owned).bytes;
    return LSTC_OK;
}

extern "C" lstc_status lstc_count(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows,
                                  uint32_t n_owned, uint64_t m, uint64_t *d_count, uint64_t *d_per_row,
                                  const lstc_tune *tune, void *d_ws, size_t ws_bytes,
                                  lstc_stream s) {
    (void)tune;
    hipStream_t st = (hipStream_t)s;
    if (!d_off || !d_count) return fail(LSTC_ERR_INVALID, "lstc_count: NULL offsets or output");
    if (n_owned > n_rows) return fail(LSTC_ERR_INVALID, "lstc_count: n_owned %u > n_rows %u", n_owned, n_rows);
    if (n_rows == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_count: n_rows must be < 2^32-1");
    if (m >= 0xFFFFF000ull) return fail(LSTC_ERR_INVALID, "lstc_count: m = %llu arcs exceeds 2^32-4096",
                                        (unsigned long long)m);
    CountLayout L = count_layout(d_ws, ws_bytes, n_owned);
    if (!d_ws || ws_bytes < L.bytes)
        return fail(LSTC_ERR_WORKSPACE, "lstc_count: workspace %zu < %zu bytes", ws_bytes, L.bytes);
    unsigned long long *out = (unsigned long long *)d_count;
    hipLaunchKernelGGL(k_count_init, dim3(1), dim3(64), 0, st, L.bin_counts, out);
    LSTC_LAUNCHED("k_count_init");
    if (d_per_row) LSTC_TRY(fill_u64(d_per_row, 0, n_owned, st));
    if (n_owned == 0) return LSTC_OK;
    if (!d_col) return fail(LSTC_ERR_INVALID, "lstc_count: NULL columns");

    // d+ bins sized to the LDS table of each kernel instance (load factor <= 1/4)
    BinBounds bb = {{2, 17, 65, 257, 0xFFFFFFFFu}};
    const int cus = device_cus();
    {
        uint32_t grid = hmin<uint32_t>(ceil_div_u32(n_owned, 256), (uint32_t)cus * 16);
        hipLaunchKernelGGL(k_bin_rows, dim3(grid), dim3(256), 0, st, d_off, n_owned, bb, L.bins,
                           n_owned, L.bin_counts);
        LSTC_LAUNCHED("k_bin_rows");
    }
    unsigned long long *pr = (unsigned long long *)d_per_row;
    const uint32_t *bins = L.bins;
    const size_t cap = n_owned;
    hipLaunchKernelGGL((k_count_rows<12, 4>), dim3(cus * 2), dim3(256), 0, st, d_off, d_col,
                       bins + 3 * cap, L.bin_counts + 3, out, pr, 1024u, n_rows);
    LSTC_LAUNCHED("k_count_rows<12>");
    hipLaunchKernelGGL((k_count_rows<10, 4>), dim3(cus * 6), dim3(256), 0, st, d_off, d_col,
                       bins + 2 * cap, L.bin_counts + 2, out, pr, 256u, n_rows);
    LSTC_LAUNCHED("k_count_rows<10>");
    hipLaunchKernelGGL((k_count_rows<8, 4>), dim3(cus * 8), dim3(256), 0, st, d_off, d_col,
                       bins + 1 * cap, L.bin_counts + 1, out, pr, 64u, n_rows);
    LSTC_LAUNCHED("k_count_rows<8>");
    hipLaunchKernelGGL((k_count_rows<6, 4>), dim3(cus * 8), dim3(256), 0, st, d_off, d_col,
                       bins + 0 * cap, L.bin_counts + 0, out, pr, 16u, n_rows);
    LSTC_LAUNCHED("k_count_rows<6>");
    return LSTC_OK;
}
