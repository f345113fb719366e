// inlists.hip -- target-side in-lists of an oriented CSR (enables split counting
// for any caller CSR, e.g. the halo subgraphs of a partition).
//
// in-list(w) = { u < n_owned : u -> w is an arc and d+(u) <= d+(w) }, ascending,
// where d+ is the row length in this CSR.  Together with the kind A filter
// d+(w) < d+(u) this assigns every arc leaving an owned row to exactly one side:
// the count over owned rows is unchanged, and each arc streams its shorter list.
//
// Steps: per row, count and emit keys (u << 32 | w) of the kept arcs in row
// order; a stable LSD radix sort on the low (target) bits groups them by
// target with sources still ascending; in_col takes the sources and in_off[w]
// is found by binary search over the sorted targets.
//
// Segment words (lstc_segments, or built in the same passes by
// lstc_in_lists_segments): for every entry v of col[] and in_col[], the word
// start(N+(v)) | d+(v) << 32, so the counting walk reads one coalesced word per
// segment instead of off[v] and off[v+1] at a random place.
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

// The kept arcs (u -> w with u < n_owned, w a row and d+(u) <= d+(w)) of a
// block's kInlRows consecutive rows of [0, n_walk) (the owned rows), walked flat in row order,
// 256 arcs at a time: each arc finds its row by binary search of the rows'
// offsets in LDS, so loads of col are coalesced however long the rows are.
// Pass 0 counts the kept arcs per row and, with SEG, writes the segment word of
// every arc it walks (it gathers the target's row offsets anyway); pass 1 writes their
// keys u << 32 | w in arc order from pos[first row of the block] on (the kept
// arcs of consecutive rows are consecutive), with SEG taking each target's
// length from the word pass 0 wrote (coalesced) instead of gathering it again.
static constexpr uint32_t kInlRows = 256;

template <bool EMIT, bool SEG>
__global__ __launch_bounds__(256) void k_inl_flat(const uint64_t *__restrict__ off, const uint32_t *__restrict__ col,
                                                  uint32_t n_rows, uint32_t n_owned, uint32_t n_walk,
                                                  uint32_t *__restrict__ cnt, const uint64_t *__restrict__ pos,
                                                  uint64_t *__restrict__ keys, uint64_t *__restrict__ col_seg) {
    __shared__ uint64_t s_off[kInlRows + 1];
    __shared__ uint32_t s_cnt[kInlRows];
    __shared__ uint32_t s_w[4];
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (!EMIT && blockIdx.x == 0 && t == 0) cnt[n_walk] = 0;
    for (uint64_t r0 = (uint64_t)blockIdx.x * kInlRows; r0 < n_walk; r0 += (uint64_t)gridDim.x * kInlRows) {
        const uint32_t R = n_walk - r0 < kInlRows ? (uint32_t)(n_walk - r0) : kInlRows;
        if (t <= R) s_off[t] = off[r0 + t];
        if (t == 0 && R == kInlRows) s_off[kInlRows] = off[r0 + kInlRows];
        if (!EMIT) s_cnt[t] = 0;
        __syncthreads();
        const uint64_t a0 = s_off[0], a1 = s_off[R];
        uint64_t outp = EMIT ? pos[r0] : 0;
        for (uint64_t c0 = a0; c0 < a1; c0 += 256) {
            const uint64_t i = c0 + t;
            bool keep = false;
            uint32_t row = 0, w = 0;
            if (i < a1) {
                uint32_t lo = 0, hi = R;  // s_off[lo] <= i < s_off[hi]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_off[mid] <= i) lo = mid; else hi = mid;
                }
                row = lo;
                w = col[i];
                uint64_t dw = 0;
                if (SEG && EMIT) {
                    dw = col_seg[i] >> 32;  // 0 for ids that are no row
                } else if (w < n_rows) {
                    const uint64_t ow = off[w];
                    dw = off[w + 1] - ow;
                    if (SEG) col_seg[i] = (uint64_t)(uint32_t)ow | (dw << 32);
                } else if (SEG) {
                    col_seg[i] = 0;
                }
                keep = r0 + row < n_owned && w < n_rows && s_off[row + 1] - s_off[row] <= dw;
            }
            if (!EMIT) {
                if (keep) atomicAdd(&s_cnt[row], 1u);
            } else {
                const uint64_t mk = ballot(keep);
                if (lane == 0) s_w[wv] = (uint32_t)__popcll(mk);
                __syncthreads();
                uint32_t before = 0;
                for (uint32_t x = 0; x < wv; ++x) before += s_w[x];
                const uint32_t total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
                if (keep) keys[outp + before + mask_rank(mk)] = ((r0 + row) << 32) | w;
                outp += total;
                __syncthreads();
            }
        }
        if (!EMIT) {
            __syncthreads();
            if (t < R) cnt[r0 + t] = s_cnt[t];
        }
        __syncthreads();
    }
}

// in_col (and with in_seg their segment words) from the sorted keys.
__global__ void k_inl_cols(const uint64_t *__restrict__ keys, uint64_t k, const uint64_t *__restrict__ off,
                           uint32_t *__restrict__ in_col, uint64_t *__restrict__ in_seg) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = (uint32_t)(keys[i] >> 32);  // a row below n_owned
        in_col[i] = u;
        if (in_seg) {
            const uint64_t b = off[u];
            in_seg[i] = (uint64_t)(uint32_t)b | ((off[u + 1] - b) << 32);
        }
    }
}

// in_off[w] = number of sorted keys whose target is below w
__global__ void k_inl_offsets(const uint64_t *__restrict__ keys, uint64_t k, uint32_t n_rows,
                              uint64_t *__restrict__ in_off) {
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w <= n_rows;
         w += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = k;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if ((uint32_t)keys[mid] < w) lo = mid + 1; else hi = mid;
        }
        in_off[w] = lo;
    }
}

__global__ void k_seg_words(const uint32_t *__restrict__ ids, uint64_t n, const uint64_t *__restrict__ off,
                            uint32_t n_rows, uint64_t *__restrict__ words) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // 8 entries per thread and step: their ids, then their row offsets, are loaded together
    for (uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i0 < n; i0 += 8 * stride) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint64_t i = i0 + j * stride;
            v[j] = i < n ? ids[i] : 0xFFFFFFFFu;
        }
        uint64_t b[8], e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool ok = v[j] < n_rows;  // ids >= n_rows (no list): length 0
            b[j] = ok ? off[v[j]] : 0;
            e[j] = ok ? off[v[j] + 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint64_t i = i0 + j * stride;
            if (i >= n) break;
            words[i] = (uint64_t)(uint32_t)b[j] | ((e[j] - b[j]) << 32);
        }
    }
}

static int bits_for(uint32_t n) {
    int b = 8;
    while (b < 32 && ((uint64_t)n >> b) != 0) b += 8;
    return b;
}

static size_t inl_ws(uint32_t n_rows, uint64_t m, size_t *parts) {
    Carver c(nullptr, 0);
    c.take<uint32_t>((size_t)n_rows + 1);  // kept arcs per row
    c.take<uint64_t>((size_t)n_rows + 1);  // their exclusive prefix
    c.take<uint64_t>(m ? m : 1);           // keys
    c.take<uint64_t>(m ? m : 1);           // sort buffer
    const size_t head = c.need();
    const size_t tail = hmax(scan_ws_bytes((uint64_t)n_rows + 1), radix_sort_ws_bytes(m));
    if (parts) *parts = head;
    return head + tail + 256;
}

// In-lists, and with col_seg / in_seg (both or neither) the segment words of
// the col entries of owned rows (the ones the split count reads) and of every
// in-list entry, in the same passes.
static lstc_status in_lists_impl(const char *who, const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows,
                                 uint32_t n_owned, uint64_t m, uint64_t *d_in_off, uint32_t *d_in_col,
                                 uint64_t *d_col_seg, uint64_t *d_in_seg, uint64_t *n_entries, void *d_ws,
                                 size_t ws_bytes, hipStream_t st) {
    if (n_owned > n_rows) return fail(LSTC_ERR_INVALID, "%s: n_owned %u > n_rows %u", who, n_owned, n_rows);
    if (n_rows == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "%s: n_rows must be < 2^32-1", who);
    if (!d_off || !d_in_off || (m && (!d_col || !d_in_col))) return fail(LSTC_ERR_INVALID, "%s: NULL array", who);
    size_t head = 0;
    const size_t need = inl_ws(n_rows, m, &head);
    if (!d_ws || ws_bytes < need) return fail(LSTC_ERR_WORKSPACE, "%s: workspace %zu < %zu", who, ws_bytes, need);
    Carver c(d_ws, ws_bytes);
    uint32_t *cnt = c.take<uint32_t>((size_t)n_rows + 1);
    uint64_t *pos = c.take<uint64_t>((size_t)n_rows + 1);
    uint64_t *keys = c.take<uint64_t>(m ? m : 1);
    uint64_t *alt = c.take<uint64_t>(m ? m : 1);
    char *tail = (char *)d_ws + head;
    const size_t tail_bytes = ws_bytes - head;
    const bool seg = d_col_seg != nullptr;
    const uint32_t n_walk = n_owned;  // the kept arcs and the segment words the count reads: owned rows
    const uint32_t grid_rows =
        (uint32_t)hmin<uint64_t>(((uint64_t)n_walk + kInlRows) / kInlRows, (uint64_t)device_cus() * 16);
    if (seg)
        hipLaunchKernelGGL((k_inl_flat<false, true>), dim3(grid_rows), dim3(256), 0, st, d_off, d_col, n_rows, n_owned,
                           n_walk, cnt, (const uint64_t *)nullptr, (uint64_t *)nullptr, d_col_seg);
    else
        hipLaunchKernelGGL((k_inl_flat<false, false>), dim3(grid_rows), dim3(256), 0, st, d_off, d_col, n_rows,
                           n_owned, n_walk, cnt, (const uint64_t *)nullptr, (uint64_t *)nullptr, (uint64_t *)nullptr);
    LSTC_LAUNCHED("k_inl_flat<count>");
    LSTC_TRY(scan_exclusive_u32_to_u64(cnt, pos, (uint64_t)n_walk + 1, nullptr, tail, tail_bytes, st));
    uint64_t k = 0;
    LSTC_HIP(hipMemcpyAsync(&k, pos + n_walk, sizeof(k), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    if (k > m) return fail(LSTC_ERR_INVALID, "%s: %llu entries exceed m", who, (unsigned long long)k);
    if (n_owned) {  // keys come from owned rows only
        const uint32_t grid_own =
            (uint32_t)hmin<uint64_t>(((uint64_t)n_owned + kInlRows) / kInlRows, (uint64_t)device_cus() * 16);
        if (seg)
            hipLaunchKernelGGL((k_inl_flat<true, true>), dim3(grid_own), dim3(256), 0, st, d_off, d_col, n_rows,
                               n_owned, n_owned, (uint32_t *)nullptr, (const uint64_t *)pos, keys, d_col_seg);
        else
            hipLaunchKernelGGL((k_inl_flat<true, false>), dim3(grid_own), dim3(256), 0, st, d_off, d_col, n_rows,
                               n_owned, n_owned, (uint32_t *)nullptr, (const uint64_t *)pos, keys, (uint64_t *)nullptr);
        LSTC_LAUNCHED("k_inl_flat<keys>");
    }
    uint64_t *sorted = keys;
    LSTC_TRY(radix_sort_u64(keys, alt, k, bits_for(n_rows), &sorted, tail, tail_bytes, st));
    const uint32_t grid_k = (uint32_t)hmin<uint64_t>((k + 255) / 256 + 1, 65535);
    if (k) {
        hipLaunchKernelGGL(k_inl_cols, dim3(grid_k), dim3(256), 0, st, sorted, k, d_off, d_in_col, d_in_seg);
        LSTC_LAUNCHED("k_inl_cols");
    }
    const uint32_t grid_w = (uint32_t)hmin<uint64_t>(((uint64_t)n_rows + 256) / 256, 65535);
    hipLaunchKernelGGL(k_inl_offsets, dim3(grid_w), dim3(256), 0, st, sorted, k, n_rows, d_in_off);
    LSTC_LAUNCHED("k_inl_offsets");
    LSTC_HIP(hipStreamSynchronize(st));
    if (n_entries) *n_entries = k;
    return LSTC_OK;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_in_lists_workspace(uint32_t n_rows, uint32_t n_owned, uint64_t m, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_in_lists_workspace: NULL bytes");
    if (n_owned > n_rows) return fail(LSTC_ERR_INVALID, "lstc_in_lists_workspace: n_owned > n_rows");
    *bytes = inl_ws(n_rows, m, nullptr);
    return LSTC_OK;
}

extern "C" lstc_status lstc_in_lists(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint32_t n_owned,
                                     uint64_t m, uint64_t *d_in_off, uint32_t *d_in_col, uint64_t *n_entries,
                                     void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_in_lists");
    return in_lists_impl("lstc_in_lists", d_off, d_col, n_rows, n_owned, m, d_in_off, d_in_col, nullptr, nullptr,
                         n_entries, d_ws, ws_bytes, (hipStream_t)s);
}

extern "C" lstc_status lstc_in_lists_segments(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows,
                                              uint32_t n_owned, uint64_t m, uint64_t *d_in_off, uint32_t *d_in_col,
                                              uint64_t *d_col_seg, uint64_t *d_in_seg, uint64_t *n_entries,
                                              void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_in_lists_segments");
    if (m >= 0xFFFFF000ull)
        return fail(LSTC_ERR_INVALID, "lstc_in_lists_segments: m = %llu exceeds 2^32-4096", (unsigned long long)m);
    if (m && (!d_col_seg || !d_in_seg)) return fail(LSTC_ERR_INVALID, "lstc_in_lists_segments: NULL segment words");
    return in_lists_impl("lstc_in_lists_segments", d_off, d_col, n_rows, n_owned, m, d_in_off, d_in_col,
                         m ? d_col_seg : nullptr, m ? d_in_seg : nullptr, n_entries, d_ws, ws_bytes, (hipStream_t)s);
}

extern "C" lstc_status lstc_segments(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint64_t m,
                                     const uint32_t *d_in_col, uint64_t n_in, uint64_t *d_col_seg, uint64_t *d_in_seg,
                                     lstc_stream s) {
    TraceRange range_("lstc_segments");
    hipStream_t st = (hipStream_t)s;
    if (n_rows == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_segments: n_rows must be < 2^32-1");
    if (m >= 0xFFFFF000ull)
        return fail(LSTC_ERR_INVALID, "lstc_segments: m = %llu exceeds 2^32-4096", (unsigned long long)m);
    if (n_in > m)
        return fail(LSTC_ERR_INVALID, "lstc_segments: %llu in-list entries exceed m", (unsigned long long)n_in);
    if (!d_off || (m && (!d_col || !d_col_seg)) || (n_in && (!d_in_col || !d_in_seg)))
        return fail(LSTC_ERR_INVALID, "lstc_segments: NULL array");
    if (m) {
        const uint32_t grid = (uint32_t)hmin<uint64_t>((m + 2047) / 2048, (uint64_t)device_cus() * 32);
        hipLaunchKernelGGL(k_seg_words, dim3(grid), dim3(256), 0, st, d_col, m, d_off, n_rows, d_col_seg);
        LSTC_LAUNCHED("k_seg_words");
    }
    if (n_in) {
        const uint32_t grid = (uint32_t)hmin<uint64_t>((n_in + 2047) / 2048, (uint64_t)device_cus() * 32);
        hipLaunchKernelGGL(k_seg_words, dim3(grid), dim3(256), 0, st, d_in_col, n_in, d_off, n_rows, d_in_seg);
        LSTC_LAUNCHED("k_seg_words");
    }
    return LSTC_OK;
}
