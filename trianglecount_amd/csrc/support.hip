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
//   one wave per owned row u (and per tile of 1024 keys of N+(u), for longer
//   rows): the tile is staged in LDS, sorted as the CSR keeps it, with one
//   counter per key (the arcs u -> x);
//   the lists N+(w), w in N+(u), are walked flat, 64 segments per window and 64
//   entries per round; each lane finds its segment by a binary search of the
//   window's prefix sums and its entry x by a binary search of the tile;
//   a hit adds to the tile counter of x (LDS), to the window counter of w (LDS),
//   and, with global atomics, to the support of arc w -> x and the count of x;
//   window and tile counters go out with one atomic each.
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint32_t kSupTile = 1024;  // keys of N+(u) per tile
static constexpr int kSupWaves = 4;

// Last index i in [0, n) with a[i] <= t (a ascending, a[0] <= t).
__device__ __forceinline__ uint32_t last_le(const uint32_t *a, uint32_t n, uint32_t t) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kSupWaves * 64) void k_support(const uint64_t *__restrict__ off,
                                                            const uint32_t *__restrict__ col, uint32_t n_rows,
                                                            uint32_t n_owned, uint32_t *__restrict__ support,
                                                            unsigned long long *__restrict__ per_vertex) {
    __shared__ uint32_t s_key[kSupWaves][kSupTile];
    __shared__ uint32_t s_cnt[kSupWaves][kSupTile];
    __shared__ uint32_t s_start[kSupWaves][64];  // window: exclusive prefix of segment lengths
    __shared__ uint64_t s_base[kSupWaves][64];   //          column index of each segment's first entry
    __shared__ uint32_t s_hits[kSupWaves][64];   //          triangles found per segment
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t *key = s_key[wv], *cnt = s_cnt[wv];
    const uint64_t nw = (uint64_t)gridDim.x * kSupWaves;
    for (uint64_t u = (uint64_t)blockIdx.x * kSupWaves + wv; u < n_owned; u += nw) {
        const uint64_t rb = off[u];
        const uint32_t K = (uint32_t)(off[u + 1] - rb);
        if (K < 2) continue;
        unsigned long long row_tri = 0;
        for (uint32_t t0 = 0; t0 < K; t0 += kSupTile) {
            const uint32_t T = min(K - t0, kSupTile);
            for (uint32_t i = lane; i < T; i += 64) {
                key[i] = col[rb + t0 + i];
                cnt[i] = 0;
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
                const uint32_t incl = wave_incl_scan(len, lane);
                const uint32_t total = read_lane(incl, 63);
                s_start[wv][lane] = incl - len;
                s_base[wv][lane] = wb;
                s_hits[wv][lane] = 0;
                const uint32_t nseg = min(64u, K - j0);
                for (uint32_t B = 0; B < total; B += 64) {
                    const uint32_t t = B + lane;
                    if (t < total) {
                        // the segment holding flat position t: the last one starting at or before t
                        // (empty segments share their start with the next one, so they are passed over)
                        const uint32_t sg = last_le(s_start[wv], nseg, t);
                        const uint64_t pos = s_base[wv][sg] + (t - s_start[wv][sg]);
                        const uint32_t x = col[pos];
                        if (x >= lo && x <= hi) {
                            const uint32_t i = last_le(key, T, x);
                            if (key[i] == x) {
                                atomicAdd(&cnt[i], 1u);            // arc u -> x
                                atomicAdd(&s_hits[wv][sg], 1u);   // arc u -> w
                                if (support) atomicAdd(&support[pos], 1u);  // arc w -> x
                                if (per_vertex) atomicAdd(&per_vertex[x], 1ull);
                            }
                        }
                    }
                }
                if (lane < nseg) {
                    const uint32_t h = s_hits[wv][lane];
                    if (h) {
                        if (support) atomicAdd(&support[rb + j0 + lane], h);
                        if (per_vertex) atomicAdd(&per_vertex[col[rb + j0 + lane]], (unsigned long long)h);
                        row_tri += h;
                    }
                }
            }
            if (support)
                for (uint32_t i = lane; i < T; i += 64)
                    if (cnt[i]) atomicAdd(&support[rb + t0 + i], cnt[i]);
        }
        const unsigned long long tot = wave_sum64(row_tri);
        if (per_vertex && lane == 0 && tot) atomicAdd(&per_vertex[u], tot);
    }
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_triangle_support(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows,
                                             uint32_t n_owned, uint64_t m, uint32_t *d_support,
                                             uint64_t *d_per_vertex, lstc_stream s) {
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
