// count.hip -- lstc_count / lstc_count_split: exact triangle count of one oriented CSR.
//
// Semantics follow Find_Triangle + cub::DeviceReduce::Sum of
// src/WithoutPartition/ShareKernelTCV62_update_v1.cu:43-191,264,271-277 and the
// owned-row launch of src/WithPartition/without_sorting_tc.cu:57-219,400,422-428.
// The work is done by the persistent engine of count_engine.hip; a single graph
// travels as a kernel argument, so the call is stream-ordered and capturable.
#include "count_engine.h"

using namespace lstc;

extern "C" lstc_status lstc_count_workspace(uint32_t n_rows, uint32_t n_owned, uint64_t m,
                                            const lstc_tune *tune, size_t *bytes) {
    (void)n_owned; (void)tune;
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_count_workspace: bytes is NULL");
    *bytes = count_ws_bytes(1, n_rows, m);  // split counting plans every row
    return LSTC_OK;
}

extern "C" lstc_status lstc_count(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows,
                                  uint32_t n_owned, uint64_t m, uint64_t *d_count, uint64_t *d_per_row,
                                  const lstc_tune *tune, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_count");
    if (!d_count) return fail(LSTC_ERR_INVALID, "lstc_count: NULL output");
    HostGraph h{d_off, d_col, d_per_row, n_rows, n_owned, m};
    LSTC_TRY(validate_graph(h, "lstc_count"));
    return count_launch(&h, 1, nullptr, nullptr, (unsigned long long *)d_count, tune, d_ws, ws_bytes,
                        (hipStream_t)s);
}

extern "C" lstc_status lstc_count_split(const uint64_t *d_off, const uint32_t *d_col, const uint64_t *d_in_off,
                                        const uint32_t *d_in_col, const uint64_t *d_col_seg,
                                        const uint64_t *d_in_seg, uint32_t n_rows, uint32_t n_owned, uint64_t m,
                                        uint64_t *d_count, const lstc_tune *tune, void *d_ws, size_t ws_bytes,
                                        lstc_stream s) {
    TraceRange range_("lstc_count_split");
    if (!d_count) return fail(LSTC_ERR_INVALID, "lstc_count_split: NULL output");
    if (!d_in_off || !d_in_col) return fail(LSTC_ERR_INVALID, "lstc_count_split: NULL in-lists");
    HostGraph h{d_off, d_col, nullptr, n_rows, n_owned, m, d_in_off, d_in_col, d_col_seg, d_in_seg};
    LSTC_TRY(validate_graph(h, "lstc_count_split"));
    return count_launch(&h, 1, nullptr, nullptr, (unsigned long long *)d_count, tune, d_ws, ws_bytes,
                        (hipStream_t)s);
}

extern "C" lstc_status lstc_count_split_stream(const uint64_t *d_off, const uint32_t *d_col, const uint64_t *d_in_off,
                                               const uint32_t *d_in_col, const uint64_t *d_col_seg,
                                               const uint64_t *d_in_seg, const uint32_t *d_scol, uint64_t scol_len,
                                               uint32_t n_rows, uint32_t n_owned, uint64_t m, uint64_t *d_count,
                                               const lstc_tune *tune, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_count_split_stream");
    if (!d_count) return fail(LSTC_ERR_INVALID, "lstc_count_split_stream: NULL output");
    if (!d_in_off || !d_in_col || !d_col_seg || !d_in_seg || (m && !d_scol))
        return fail(LSTC_ERR_INVALID, "lstc_count_split_stream: NULL in-lists, segment words or streaming copy");
    HostGraph h{d_off, d_col, nullptr, n_rows, n_owned, m, d_in_off, d_in_col, d_col_seg, d_in_seg, d_scol, scol_len};
    LSTC_TRY(validate_graph(h, "lstc_count_split_stream"));
    return count_launch(&h, 1, nullptr, nullptr, (unsigned long long *)d_count, tune, d_ws, ws_bytes,
                        (hipStream_t)s);
}
