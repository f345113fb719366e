// streamcopy.hip -- a row-aligned streaming copy of any CSR prepared for the
// split count (lstc_stream_copy).
//
// The split count streams, for every segment word, the list N+(w) the word
// names (start | d+(w) << 32).  Read from col, such a list usually begins in
// the middle of a 128-byte line; the copy scol holds every row again starting
// on a 32-entry (128-byte) boundary, and the words are rewritten to point into
// it, so each streamed list begins on a line.  lstc_clean_orient_stream builds
// the same layout during the preprocessing; this call builds it for a CSR that
// got its in-lists and words elsewhere (lstc_in_lists_segments: halo parts, a
// CSR read from a file).  No counterpart in the reference, which reads col.
//
//   pad    : plen[v] = d+(v) rounded up to 32; astart = exclusive scan
//   copy   : scol[astart[v] + k] = col[off[v] + k], eight lanes per row (a
//            wave copies eight consecutive rows, which lie next to each other
//            in col; most rows are short, so wider groups would idle)
//   words  : a word with a list (d+ != 0) of entry i gets start astart[id[i]],
//            gathered from a 32-bit copy of astart (every start is below 2^32);
//            the words and ids are read once, with the non-temporal hint, so
//            the gathered table keeps the cache
#include "lstc_internal.h"

namespace lstc {

static constexpr uint32_t kScLanes = 8;  // lanes per row in the copy

__global__ void k_sc_pad(const uint64_t *__restrict__ off, uint32_t n_rows, uint32_t *__restrict__ plen) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n_rows; v += (uint64_t)gridDim.x * blockDim.x)
        plen[v] = ((uint32_t)(off[v + 1] - off[v]) + 31u) & ~31u;
}

__global__ void k_sc_narrow(const uint64_t *__restrict__ a, uint32_t n, uint32_t *__restrict__ a32) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        a32[v] = (uint32_t)a[v];
}

// astart[v] + d(v) <= astart[v + 1] <= the copy's capacity (the rows are padded to 32).
__global__ __launch_bounds__(256) void k_sc_copy(const uint64_t *__restrict__ off, const uint32_t *__restrict__ col,
                                                 uint32_t n_rows, const uint32_t *__restrict__ astart,
                                                 uint32_t *__restrict__ scol) {
    const uint32_t gl = threadIdx.x % kScLanes, groups = blockDim.x / kScLanes;
    for (uint64_t v = (uint64_t)blockIdx.x * groups + threadIdx.x / kScLanes; v < n_rows;
         v += (uint64_t)gridDim.x * groups) {
        const uint64_t o = off[v], d = off[v + 1] - o, a = astart[v];  // a < 2^32
        for (uint64_t k = gl; k < d; k += kScLanes) scol[a + k] = col[o + k];
    }
}

__global__ void k_sc_words(uint64_t *__restrict__ words, const uint32_t *__restrict__ ids, uint64_t n,
                           const uint32_t *__restrict__ astart) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t w = __builtin_nontemporal_load(words + i);
        if (w >> 32) words[i] = (w & ~0xFFFFFFFFull) | astart[__builtin_nontemporal_load(ids + i)];
    }
}

struct ScLayout {
    uint32_t *plen;
    uint64_t *astart, *total;
    uint32_t *astart32;  // astart narrowed: every start is below 2^32
    void *scan_ws;
    size_t scan_bytes, bytes;
};

static ScLayout sc_layout(void *ws, size_t ws_bytes, uint32_t n_rows) {
    Carver c(ws, ws_bytes);
    ScLayout L;
    L.plen = c.take<uint32_t>((size_t)n_rows + 1);
    L.astart = c.take<uint64_t>((size_t)n_rows + 1);
    L.total = c.take<uint64_t>(1);
    L.astart32 = c.take<uint32_t>((size_t)n_rows + 1);
    L.scan_bytes = scan_ws_bytes((uint64_t)n_rows + 1);
    L.scan_ws = c.take<char>(L.scan_bytes);
    L.bytes = c.need();
    return L;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_stream_copy_workspace(uint32_t n_rows, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_stream_copy_workspace: NULL bytes");
    *bytes = sc_layout(nullptr, 0, n_rows).bytes;
    return LSTC_OK;
}

extern "C" lstc_status lstc_stream_copy(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint64_t m,
                                        uint64_t n_col_words, uint64_t *d_col_seg, const uint32_t *d_in_col,
                                        uint64_t n_in, uint64_t *d_in_seg, uint32_t *d_scol, uint64_t scol_cap,
                                        uint64_t *scol_out, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_stream_copy");
    const char *who = "lstc_stream_copy";
    if (!scol_out) return fail(LSTC_ERR_INVALID, "%s: NULL size output", who);
    if (n_rows == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "%s: n_rows must be < 2^32-1", who);
    const uint64_t need = m + 31ull * n_rows;
    if (need >= (1ull << 32)) return fail(LSTC_ERR_INVALID, "%s: m + 31 n_rows = %llu must be < 2^32", who,
                                          (unsigned long long)need);
    if (scol_cap < need)
        return fail(LSTC_ERR_INVALID, "%s: copy capacity %llu < m + 31 n_rows = %llu", who,
                    (unsigned long long)scol_cap, (unsigned long long)need);
    if (n_col_words > m) return fail(LSTC_ERR_INVALID, "%s: %llu column words for %llu arcs", who,
                                     (unsigned long long)n_col_words, (unsigned long long)m);
    if ((n_rows && !d_off) || (m && (!d_col || !d_scol)) || (n_col_words && !d_col_seg) ||
        (n_in && (!d_in_col || !d_in_seg)))
        return fail(LSTC_ERR_INVALID, "%s: NULL arrays", who);
    ScLayout L = sc_layout(d_ws, ws_bytes, n_rows);
    if (!d_ws || ws_bytes < L.bytes)
        return fail(LSTC_ERR_WORKSPACE, "%s: workspace %zu < %zu bytes", who, ws_bytes, L.bytes);
    hipStream_t st = (hipStream_t)s;
    const int cus = device_cus();
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n_rows + 255) / 256 + 1, (uint64_t)cus * 16);
    hipLaunchKernelGGL(k_sc_pad, dim3(grid_n), dim3(256), 0, st, d_off, n_rows, L.plen);
    LSTC_LAUNCHED("k_sc_pad");
    LSTC_TRY(scan_exclusive_u32_to_u64(L.plen, L.astart, n_rows, L.total, L.scan_ws, L.scan_bytes, st));
    if (n_rows) {
        hipLaunchKernelGGL(k_sc_narrow, dim3(grid_n), dim3(256), 0, st, (const uint64_t *)L.astart, n_rows, L.astart32);
        LSTC_LAUNCHED("k_sc_narrow");
    }
    if (m) {
        const uint64_t per_block = 256 / kScLanes;
        const uint32_t grid_c = (uint32_t)hmin<uint64_t>(((uint64_t)n_rows + per_block - 1) / per_block, (uint64_t)cus * 16);
        hipLaunchKernelGGL(k_sc_copy, dim3(grid_c), dim3(256), 0, st, d_off, d_col, n_rows, (const uint32_t *)L.astart32,
                           d_scol);
        LSTC_LAUNCHED("k_sc_copy");
    }
    const uint32_t grid_w = (uint32_t)hmin<uint64_t>((hmax(n_col_words, n_in) + 255) / 256 + 1, (uint64_t)cus * 16);
    if (n_col_words) {
        hipLaunchKernelGGL(k_sc_words, dim3(grid_w), dim3(256), 0, st, d_col_seg, d_col, n_col_words,
                           (const uint32_t *)L.astart32);
        LSTC_LAUNCHED("k_sc_words(col)");
    }
    if (n_in) {
        hipLaunchKernelGGL(k_sc_words, dim3(grid_w), dim3(256), 0, st, d_in_seg, d_in_col, n_in,
                           (const uint32_t *)L.astart32);
        LSTC_LAUNCHED("k_sc_words(in)");
    }
    LSTC_HIP(hipMemcpyAsync(scol_out, L.total, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    return LSTC_OK;
}
