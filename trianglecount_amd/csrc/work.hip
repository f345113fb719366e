// work.hip -- what one count moves, and how long its kernel runs.
//
// lstc_work_stats: the work of the count engine's plan for a graph, counted on
// the device from the graph itself (items, table-fill entries, located lists,
// streamed list entries) and turned into the engine's algorithmic bytes: the
// bytes the kernels read and write if every list they touch came from HBM once
// per use.  This is the figure a roofline of the kernel as designed uses; the
// merge model of the reference's kernel (both lists of every arc, as
// src/WithoutPartition/ShareKernelTCV62_update_v1.cu:43-191 walks them) is a
// different, larger figure.
//
// lstc_timing_*: per-launch kernel time of the count kernel from hipEvents
// recorded on the launch stream, the way the reference times its kernel with
// cudaEvents (ShareKernelTCV62_update_v1.cu:263-290).
#include <atomic>
#include <mutex>
#include <utility>
#include <vector>

#include "count_engine.h"
#include "wave.h"

namespace lstc {

enum { W_ROWS, W_ITEMS, W_CHUNKS, W_TINY, W_HASHED, W_SEGS, W_TINY_SEGS, W_STREAMED, W_N };

// One wave per planned row; the terms mirror what k_plan_rows emits and what
// k_count reads for each item (count_engine.hip).
__global__ __launch_bounds__(256) void k_work_stats(GraphDev G, uint32_t tiny_max, unsigned long long *__restrict__ acc) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    const bool split = G.in_off != nullptr;
    const uint64_t rows = split ? G.n_rows : G.n_owned;
    unsigned long long c[W_N];
#pragma unroll
    for (int i = 0; i < W_N; ++i) c[i] = 0;
    for (uint64_t r = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < rows; r += nw) {
        if (lane == 0) c[W_ROWS] += 1;
        const uint64_t rb = G.off[r], K = G.off[r + 1] - rb;
        if (K < 2) continue;  // no pair of out-neighbours: the planner skips the row
        const bool owned = r < G.n_owned;
        const uint64_t ib = split ? G.in_off[r] : 0, kb = split ? G.in_off[r + 1] - ib : 0;
        if (!owned && kb == 0) continue;
        const bool tiny = split && K <= tiny_max;
        if (tiny) {
            if (lane == 0) {
                c[W_TINY] += 1;
                c[W_HASHED] += K;  // the row's list, staged once
            }
        } else {
            const uint32_t sb = in_chunk(kb);
            const uint64_t nb = (kb + sb - 1) / sb, items = (owned ? 1u : 0u) + nb;
            if (lane == 0) {
                c[W_ITEMS] += items;
                c[W_CHUNKS] += nb;
                c[W_HASHED] += K * items;  // every item builds its table from N+(r)
            }
        }
        unsigned long long segs = 0;
        if (owned) {  // source side: N+(w) for w in N+(r); split keeps d+(w) < d+(r)
            for (uint64_t j = lane; j < K; j += 64) {
                const uint32_t w = G.col[rb + j];
                const uint64_t dw = w < G.n_rows ? G.off[w + 1] - G.off[w] : 0;
                if (!split || dw < K) c[W_STREAMED] += dw;
                segs += 1;
            }
        }
        for (uint64_t j = lane; j < kb; j += 64) {  // target side: N+(u) for u in r's in-list
            const uint32_t u = G.in_col[ib + j];
            c[W_STREAMED] += G.off[u + 1] - G.off[u];
            segs += 1;
        }
        if (tiny) c[W_TINY_SEGS] += segs; else c[W_SEGS] += segs;
    }
#pragma unroll
    for (int i = 0; i < W_N; ++i) {
        const unsigned long long v = wave_sum64(c[i]);
        if (lane == 0 && v) atomicAdd(acc + i, v);
    }
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------- kernel timing
static std::atomic<int> g_timing{0};
static std::mutex g_timing_mu;
static std::vector<std::pair<hipEvent_t, hipEvent_t>> g_pending, g_pool;

static bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

TimingMark timing_begin(hipStream_t st) {
    TimingMark m{nullptr, nullptr};
    if (!g_timing.load(std::memory_order_relaxed) || capturing(st)) return m;
    {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        if (!g_pool.empty()) {
            m.e0 = g_pool.back().first;
            m.e1 = g_pool.back().second;
            g_pool.pop_back();
        }
    }
    if (!m.e0 && (hipEventCreate(&m.e0) != hipSuccess || hipEventCreate(&m.e1) != hipSuccess)) return TimingMark{};
    if (hipEventRecord(m.e0, st) != hipSuccess) return TimingMark{};
    return m;
}

void timing_end(const TimingMark &m, hipStream_t st) {
    if (!m.e0) return;
    if (hipEventRecord(m.e1, st) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_pending.emplace_back(m.e0, m.e1);
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_work_stats_workspace(uint32_t G, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_work_stats_workspace: NULL bytes");
    *bytes = ((size_t)(G ? G : 1) * W_N * sizeof(unsigned long long) + 255) & ~size_t(255);
    return LSTC_OK;
}

extern "C" lstc_status lstc_work_stats(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune,
                                       lstc_work *out, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_work_stats");
    hipStream_t st = (hipStream_t)s;
    if (G == 0) return LSTC_OK;
    if (!graphs || !out) return fail(LSTC_ERR_INVALID, "lstc_work_stats: NULL argument");
    size_t need = 0;
    LSTC_TRY(lstc_work_stats_workspace(G, &need));
    if (!d_ws || ws_bytes < need) return fail(LSTC_ERR_WORKSPACE, "lstc_work_stats: workspace %zu < %zu", ws_bytes, need);
    uint64_t arcs = 0;
    for (uint32_t g = 0; g < G; ++g) arcs += graphs[g].m;
    const TuneR tr = effective_tune(tune, arcs);
    unsigned long long *acc = (unsigned long long *)d_ws;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, st, acc, G * (uint32_t)W_N);
    LSTC_LAUNCHED("k_zero_u64");
    std::vector<HostGraph> hosts(G);
    for (uint32_t g = 0; g < G; ++g) {
        const lstc_graph_desc &d = graphs[g];
        hosts[g] = HostGraph{d.off, d.col, d.per_row, d.n_rows, d.n_owned, d.m,
                              d.in_off, d.in_col, d.col_seg, d.in_seg, d.scol, d.scol_len};
        LSTC_TRY(validate_graph(hosts[g], "lstc_work_stats"));
        const GraphDev dev = to_dev(hosts[g]);
        const uint64_t rows = plan_rows(hosts[g]);
        if (rows == 0) continue;
        const uint32_t grid = (uint32_t)hmin<uint64_t>((rows + 3) / 4, (uint64_t)device_cus() * 16);
        hipLaunchKernelGGL(k_work_stats, dim3(grid), dim3(256), 0, st, dev, tr.tiny_max, acc + (size_t)g * W_N);
        LSTC_LAUNCHED("k_work_stats");
    }
    std::vector<unsigned long long> h((size_t)G * W_N);
    LSTC_HIP(hipMemcpyAsync(h.data(), acc, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    for (uint32_t g = 0; g < G; ++g) {
        const unsigned long long *c = h.data() + (size_t)g * W_N;
        const bool split = split_ok(hosts[g]);
        lstc_work w{};
        w.rows = c[W_ROWS];
        w.items = c[W_ITEMS];
        w.chunk_items = c[W_CHUNKS];
        w.tiny_rows = c[W_TINY];
        w.hashed_keys = c[W_HASHED];
        w.segments = c[W_SEGS];
        w.tiny_segments = c[W_TINY_SEGS];
        w.streamed = c[W_STREAMED];
        const uint64_t seg_bytes = seg_ok(hosts[g]) ? 8 : 20;   // a segment word, or an id and two row offsets
        const uint64_t tiny_bytes = seg_ok(hosts[g]) ? 12 : 20;  // an id and its word, or an id and two offsets
        w.bytes = 4 * (w.hashed_keys + w.streamed) + seg_bytes * w.segments + tiny_bytes * w.tiny_segments +
                  16 * (w.items + w.chunk_items) + 32 * w.tiny_rows + 16 * (w.items + w.tiny_rows) +
                  2 * w.rows * (split ? 32 : 16);
        out[g] = w;
    }
    return LSTC_OK;
}

extern "C" lstc_status lstc_timing_enable(int on) {
    g_timing.store(on ? 1 : 0);
    return LSTC_OK;
}

extern "C" lstc_status lstc_timing_read(double *kernel_ms, uint64_t *launches) {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> done;
    {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        done.swap(g_pending);
    }
    double ms = 0;
    lstc_status rc = LSTC_OK;
    for (auto &p : done) {
        float f = 0;
        hipError_t e = hipEventSynchronize(p.second);
        if (e == hipSuccess) e = hipEventElapsedTime(&f, p.first, p.second);
        if (e != hipSuccess && rc == LSTC_OK) rc = hip_fail(e, "lstc_timing_read");
        ms += f;
    }
    {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        g_pool.insert(g_pool.end(), done.begin(), done.end());
    }
    if (kernel_ms) *kernel_ms = ms;
    if (launches) *launches = done.size();
    return rc;
}
