// count_engine.h -- the persistent counting engine shared by lstc_count,
// lstc_count_batch, batch plans and the server.
#pragma once
#include "lstc_internal.h"

namespace lstc {

// A row's in-list is split into kind B items of kInChunk sources, so that one
// hub with a short list and a huge in-list (power-law graphs) becomes many
// items instead of one straggler.  Larger chunks only when the chunk index
// would not fit its 20-bit field.
static constexpr uint32_t kInChunk = 512;
static constexpr uint32_t kMaxChunks = (1u << 20) - 1;

__host__ __device__ __forceinline__ uint32_t in_chunk(uint64_t in_degree) {
    const uint64_t need = (in_degree + kMaxChunks - 1) / kMaxChunks;
    return need <= kInChunk ? kInChunk : (uint32_t)((need + 63u) & ~63ull);
}

static constexpr uint32_t kTinyMax = 8;  // largest d+ of the lane-per-arc path

// One graph as the kernels see it.
struct GraphDev {
    const uint64_t *off;
    const uint32_t *col;
    unsigned long long *per_row;  // may be null
    const uint64_t *in_off;       // target-side in-lists (split count), may be null
    const uint32_t *in_col;
    uint32_t n_rows;
    uint32_t n_owned;
    const uint64_t *col_seg;      // segment words of col[] / in_col[] (split count), may be null
    const uint64_t *in_seg;
    const uint32_t *scol;         // the columns the walk reads at segment-word starts: col, or a
                                  // copy whose rows start every 128 bytes (lstc_clean_orient_stream)
};

// The graphs of one launch, passed by value.  A single graph travels inline
// (no device table, so the launch is fully stream-ordered and capturable);
// batches use a device table and a device prefix of owned-row counts.
struct GraphSet {
    const GraphDev *table;     // device [G] or null when G == 1
    const uint64_t *row_base;  // device [G+1] or null when G == 1
    GraphDev one;
    uint64_t total_rows;       // sum of n_owned
    uint32_t G;
};

struct HostGraph {
    const uint64_t *off;
    const uint32_t *col;
    uint64_t *per_row;
    uint32_t n_rows, n_owned;
    uint64_t m;
    const uint64_t *in_off = nullptr;  // optional target-side in-lists
    const uint32_t *in_col = nullptr;
    const uint64_t *col_seg = nullptr;  // optional segment words (lstc_segments)
    const uint64_t *in_seg = nullptr;
    const uint32_t *scol = nullptr;     // optional streaming copy the segment words point into
    uint64_t scol_len = 0;              // its entries
};

// Split counting (in-lists present) applies only to whole graphs without per-row output.
// Split counting needs the in-lists and no per-row output (a row's count is
// spread over both sides).  With n_owned < n_rows the in-lists must hold only
// owned sources (lstc_in_lists does that).
inline bool split_ok(const HostGraph &h) { return h.in_off && h.in_col && !h.per_row; }
// Rows the planner visits: every row for split counting (in-list items of halo
// rows), else the owned rows.
inline uint64_t plan_rows(const HostGraph &h) { return split_ok(h) ? h.n_rows : h.n_owned; }
// Segment words are used only with the in-lists they describe.
inline bool seg_ok(const HostGraph &h) { return split_ok(h) && h.col_seg && h.in_seg; }
inline GraphDev to_dev(const HostGraph &h) {
    const bool sp = split_ok(h), sg = seg_ok(h);
    return GraphDev{h.off, h.col, (unsigned long long *)h.per_row, sp ? h.in_off : nullptr, sp ? h.in_col : nullptr,
                    h.n_rows, h.n_owned, sg ? h.col_seg : nullptr, sg ? h.in_seg : nullptr,
                    sg && h.scol ? h.scol : h.col};
}

// Resolved tuning (defaults filled in, ranges clamped).
struct TuneR {
    uint32_t table_scale;   // LDS table capacity >= table_scale * keys
    uint32_t small_max;     // rows with d+ <= small_max use a wave-private slice
    uint32_t small_group;   // small rows per ticket
    uint32_t grid_per_cu;   // persistent workgroups per CU
    uint32_t unroll;        // chunks in flight per wave (4 or 8)
    uint32_t tiny_max;      // rows with 2 <= d+ <= tiny_max (graphs with in-lists): lane-per-arc path
    uint32_t waves;         // waves per workgroup of the count kernel (2, 4 or 8)
};
TuneR resolve_tune(const lstc_tune *t);
// The tuning a launch of total_m arcs runs with: resolve_tune plus the
// small-launch default of small_max.  The autotuner reports this, so a
// returned tune replays exactly the configuration that was timed.
TuneR effective_tune(const lstc_tune *t, uint64_t total_m);

lstc_status validate_graph(const HostGraph &h, const char *who);

// Workspace for a launch over G graphs with total_rows owned rows and total_m
// arcs (excludes the batch descriptor table, which the caller places).
size_t count_ws_bytes(uint32_t G, uint64_t total_rows, uint64_t total_m);
// Largest batch one launch can address (graph ids are packed into work items).
constexpr uint32_t kMaxGraphs = 2047;

// hosts[G] describe the graphs (used for sizes, per-row zeroing and, when G == 1,
// as the inline descriptor).  For G > 1, d_table[G] and d_row_base[G+1] must
// already hold the same descriptors and the owned-row prefix on the device.
// d_counts[G] are overwritten.  Only kernels are enqueued (no copies, no
// synchronisation), so the sequence can be captured into a hipGraph.
// timed = false keeps the launch free of the optional kernel timing (which may
// create events), for callers that must not allocate.
lstc_status count_launch(const HostGraph *hosts, uint32_t G, const GraphDev *d_table, const uint64_t *d_row_base,
                         unsigned long long *d_counts, const lstc_tune *tune, void *ws, size_t ws_bytes,
                         hipStream_t st, bool timed = true);

// Kernel timing (lstc_timing_enable): when enabled and the stream is not being
// captured, timing_begin / timing_end record a hipEvent pair around a launch.
struct TimingMark {
    hipEvent_t e0, e1;
};
TimingMark timing_begin(hipStream_t st);
void timing_end(const TimingMark &m, hipStream_t st);

}  // namespace lstc
