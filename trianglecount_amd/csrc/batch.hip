// batch.hip -- many graphs in one persistent count launch (+ hipGraph replay).
//
// The reference counts the parts of a partition one after another: per part a
// host parse, H2D copy, Find_Triangle launch, synchronise, CUB reduce and D2H
// copy (src/WithPartition/without_sorting_tc.cu:244-465).  Here all graphs of a
// batch are planned into one degree-binned ticket queue and counted by one
// persistent launch of the engine (count_engine.hip); per-graph counts land in
// d_counts[G].  A plan keeps the device descriptor table and workspace
// resident and can capture the whole launch sequence into a hipGraph once, so
// a repeat costs one hipGraphLaunch.
#include <vector>

#include "count_engine.h"

namespace lstc {

static lstc_status to_hosts(const lstc_graph_desc *graphs, uint32_t G, std::vector<HostGraph> &hosts,
                            std::vector<GraphDev> &devs, std::vector<uint64_t> &base) {
    if (G == 0) return LSTC_OK;
    if (!graphs) return fail(LSTC_ERR_INVALID, "batch: NULL graph array");
    hosts.resize(G);
    devs.resize(G);
    base.resize((size_t)G + 1);
    if (G > kMaxGraphs) return fail(LSTC_ERR_INVALID, "batch: at most %u graphs per launch", kMaxGraphs);
    uint64_t acc = 0;
    for (uint32_t g = 0; g < G; ++g) {
        const lstc_graph_desc &d = graphs[g];
        hosts[g] = HostGraph{d.off, d.col, d.per_row, d.n_rows, d.n_owned, d.m,
                                     d.in_off, d.in_col, d.col_seg, d.in_seg, d.scol, d.scol_len};
        LSTC_TRY(validate_graph(hosts[g], "batch"));
        devs[g] = to_dev(hosts[g]);
        base[g] = acc;
        acc += plan_rows(hosts[g]);
    }
    base[G] = acc;
    return LSTC_OK;
}

static uint64_t planned_total(const lstc_graph_desc *graphs, uint32_t G) {
    uint64_t t = 0;
    for (uint32_t g = 0; g < G; ++g) {
        const lstc_graph_desc &d = graphs[g];
        t += plan_rows(HostGraph{d.off, d.col, d.per_row, d.n_rows, d.n_owned, d.m,
                                     d.in_off, d.in_col, d.col_seg, d.in_seg, d.scol, d.scol_len});
    }
    return t;
}

static uint64_t arcs_total(const lstc_graph_desc *graphs, uint32_t G) {
    uint64_t t = 0;
    for (uint32_t g = 0; g < G; ++g) t += graphs[g].m;
    return t;
}

static size_t table_bytes(uint32_t G) {
    Carver c(nullptr, 0);
    c.take<GraphDev>(G ? G : 1);
    c.take<uint64_t>((size_t)G + 1);
    return c.need();
}

}  // namespace lstc

using namespace lstc;

struct lstc_batch_plan {
    std::vector<HostGraph> hosts;
    uint32_t G = 0;
    lstc_tune tune{};
    bool has_tune = false;
    void *dmem = nullptr;  // descriptor table + row_base + engine workspace
    size_t dbytes = 0;
    GraphDev *d_table = nullptr;
    uint64_t *d_base = nullptr;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t *captured_counts = nullptr;
};

extern "C" lstc_status lstc_count_batch_workspace(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune,
                                                  size_t *bytes) {
    (void)tune;
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_count_batch_workspace: NULL bytes");
    if (G && !graphs) return fail(LSTC_ERR_INVALID, "lstc_count_batch_workspace: NULL graph array");
    *bytes = table_bytes(G) + count_ws_bytes(G, planned_total(graphs, G), arcs_total(graphs, G));
    return LSTC_OK;
}

extern "C" lstc_status lstc_count_batch(const lstc_graph_desc *graphs, uint32_t G, uint64_t *d_counts,
                                        const lstc_tune *tune, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_count_batch");
    hipStream_t st = (hipStream_t)s;
    std::vector<HostGraph> hosts;
    std::vector<GraphDev> devs;
    std::vector<uint64_t> base;
    LSTC_TRY(to_hosts(graphs, G, hosts, devs, base));
    if (G == 0) return LSTC_OK;
    if (!d_counts) return fail(LSTC_ERR_INVALID, "lstc_count_batch: NULL counts");
    if (G == 1)
        return count_launch(hosts.data(), 1, nullptr, nullptr, (unsigned long long *)d_counts, tune, d_ws, ws_bytes, st);
    const size_t tb = table_bytes(G);
    if (!d_ws || ws_bytes < tb) return fail(LSTC_ERR_WORKSPACE, "lstc_count_batch: workspace too small");
    Carver c(d_ws, ws_bytes);
    GraphDev *d_table = c.take<GraphDev>(G);
    uint64_t *d_base = c.take<uint64_t>((size_t)G + 1);
    LSTC_HIP(hipMemcpyAsync(d_table, devs.data(), sizeof(GraphDev) * G, hipMemcpyHostToDevice, st));
    LSTC_HIP(hipMemcpyAsync(d_base, base.data(), sizeof(uint64_t) * (G + 1), hipMemcpyHostToDevice, st));
    lstc_status rc = count_launch(hosts.data(), G, d_table, d_base, (unsigned long long *)d_counts, tune,
                                  (char *)d_ws + tb, ws_bytes - tb, st);
    // the descriptor copies read host vectors that end with this call
    hipError_t e = hipStreamSynchronize(st);
    if (rc != LSTC_OK) return rc;
    if (e != hipSuccess) return hip_fail(e, "lstc_count_batch: hipStreamSynchronize");
    return LSTC_OK;
}

extern "C" lstc_status lstc_batch_plan_create(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune,
                                              lstc_batch_plan **out) {
    if (!out) return fail(LSTC_ERR_INVALID, "lstc_batch_plan_create: NULL out");
    *out = nullptr;
    if (G == 0) return fail(LSTC_ERR_INVALID, "lstc_batch_plan_create: empty batch");
    lstc_batch_plan *p = new lstc_batch_plan();
    std::vector<GraphDev> devs;
    std::vector<uint64_t> base;
    lstc_status rc = to_hosts(graphs, G, p->hosts, devs, base);
    if (rc != LSTC_OK) { delete p; return rc; }
    p->G = G;
    if (tune) { p->tune = *tune; p->has_tune = true; }
    const size_t tb = table_bytes(G);
    p->ws_bytes = count_ws_bytes(G, base[G], arcs_total(graphs, G));
    p->dbytes = tb + p->ws_bytes;
    hipError_t e = hipMalloc(&p->dmem, p->dbytes);
    if (e != hipSuccess) { delete p; return fail(LSTC_ERR_OOM, "lstc_batch_plan_create: hipMalloc(%zu) failed", p->dbytes); }
    Carver c(p->dmem, p->dbytes);
    p->d_table = c.take<GraphDev>(G);
    p->d_base = c.take<uint64_t>((size_t)G + 1);
    p->ws = (char *)p->dmem + tb;
    e = hipMemcpy(p->d_table, devs.data(), sizeof(GraphDev) * G, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_base, base.data(), sizeof(uint64_t) * (G + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipFree(p->dmem);
        delete p;
        return hip_fail(e, "lstc_batch_plan_create");
    }
    *out = p;
    return LSTC_OK;
}

static lstc_status plan_enqueue(lstc_batch_plan *p, uint64_t *d_counts, hipStream_t st) {
    return count_launch(p->hosts.data(), p->G, p->G > 1 ? p->d_table : nullptr, p->G > 1 ? p->d_base : nullptr,
                        (unsigned long long *)d_counts, p->has_tune ? &p->tune : nullptr, p->ws, p->ws_bytes, st);
}

extern "C" lstc_status lstc_batch_plan_run(lstc_batch_plan *p, uint64_t *d_counts, lstc_stream s) {
    TraceRange range_("lstc_batch_plan_run");
    if (!p || !d_counts) return fail(LSTC_ERR_INVALID, "lstc_batch_plan_run: NULL argument");
    return plan_enqueue(p, d_counts, (hipStream_t)s);
}

extern "C" lstc_status lstc_batch_plan_launch_graph(lstc_batch_plan *p, uint64_t *d_counts, lstc_stream s) {
    TraceRange range_("lstc_batch_plan_launch_graph");
    if (!p || !d_counts) return fail(LSTC_ERR_INVALID, "lstc_batch_plan_launch_graph: NULL argument");
    if (!p->exec || p->captured_counts != d_counts) {
        // a replay of the old graph may still be queued: let it finish first
        if (p->exec) LSTC_HIP(hipDeviceSynchronize());
        if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
        if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
        LSTC_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        lstc_status rc = plan_enqueue(p, d_counts, p->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
        if (rc != LSTC_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        p->graph = g;
        LSTC_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        p->captured_counts = d_counts;
    }
    LSTC_HIP(hipGraphLaunch(p->exec, (hipStream_t)s));
    return LSTC_OK;
}

extern "C" lstc_status lstc_batch_plan_destroy(lstc_batch_plan *p) {
    if (!p) return LSTC_OK;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) {
        (void)hipStreamSynchronize(p->cap_stream);
        (void)hipStreamDestroy(p->cap_stream);
    }
    if (p->dmem) {
        (void)hipDeviceSynchronize();  // a replay may still read the descriptors
        (void)hipFree(p->dmem);
    }
    delete p;
    return LSTC_OK;
}
