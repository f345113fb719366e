// tune.hip -- autotuner for the count engine and a per-device tuning cache.
//
// The reference tunes by hand: the block size and shared tile are compile-time
// #defines (src/WithoutPartition/ShareKernelTCV62_update_v1.cu:8-9) swept with
// Nsight Compute (src/profiling/occup_{64,128,256,512}.txt).  Here the knobs of
// lstc_tune are searched at run time on the caller's own batch: coordinate
// descent from the defaults, one knob at a time over its candidate values, each
// candidate timed with hipEvents over `reps` launches after a warm-up.  Every
// candidate must reproduce the default's counts exactly, otherwise the search
// stops with LSTC_ERR_INVALID (a knob may change speed, never results).
// Winners are cached in a TSV keyed by (device, shape class).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "count_engine.h"

namespace lstc {

static void device_key(char *buf, size_t n) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
        snprintf(buf, n, "unknown");
        return;
    }
    snprintf(buf, n, "%s:%dcu", p.gcnArchName, p.multiProcessorCount);
    for (char *c = buf; *c; ++c)
        if (*c == '\t' || *c == '\n' || *c == ' ') *c = '_';
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_shape_class(const lstc_graph_desc *graphs, uint32_t G, uint32_t *shape_class) {
    if (!shape_class || (G && !graphs)) return fail(LSTC_ERR_INVALID, "lstc_shape_class: NULL argument");
    uint64_t m = 0, n = 0;
    for (uint32_t g = 0; g < G; ++g) {
        m += graphs[g].m;
        n += graphs[g].n_owned;
    }
    // class = 100 * ceil(log2 arcs) + 10 * ceil(log2 mean out-degree) + (batch ? 1 : 0)
    const uint32_t lm = m ? (uint32_t)ceil(log2((double)m)) : 0;
    const uint32_t ld = (n && m) ? (uint32_t)ceil(log2((double)m / (double)n + 1.0)) : 0;
    *shape_class = 100 * lm + 10 * (ld > 9 ? 9 : ld) + (G > 1 ? 1 : 0);
    return LSTC_OK;
}

extern "C" lstc_status lstc_tune_resolve(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune,
                                         lstc_tune *out) {
    if (!out || (G && !graphs)) return fail(LSTC_ERR_INVALID, "lstc_tune_resolve: NULL argument");
    uint64_t arcs = 0;
    for (uint32_t g = 0; g < G; ++g) arcs += graphs[g].m;
    const TuneR r = effective_tune(tune, arcs);
    lstc_tune t{};
    t.table_scale = r.table_scale;
    t.small_max = r.small_max;
    t.small_group = r.small_group;
    t.grid_per_cu = r.grid_per_cu;
    t.unroll = r.unroll;
    t.tiny_max = r.tiny_max ? r.tiny_max : 1u;
    t.waves = r.waves;
    *out = t;
    return LSTC_OK;
}

extern "C" lstc_status lstc_autotune(const lstc_graph_desc *graphs, uint32_t G, uint32_t reps, lstc_tune *best,
                                     double *best_ms, lstc_stream s) {
    TraceRange range_("lstc_autotune");
    hipStream_t st = (hipStream_t)s;
    if (!graphs || G == 0 || !best) return fail(LSTC_ERR_INVALID, "lstc_autotune: bad arguments");
    if (reps == 0) reps = 3;
    lstc_batch_plan *plan = nullptr;
    size_t need = 0;
    unsigned long long *d_counts = nullptr;
    std::vector<unsigned long long> ref(G), got(G);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    lstc_tune cur{};
    double cur_ms = 1e30;
    lstc_status rc = LSTC_OK;
    (void)need;

    // The autotuner is a set-up call: it allocates its own scratch.
    if (hipMalloc((void **)&d_counts, sizeof(unsigned long long) * G) != hipSuccess)
        return fail(LSTC_ERR_OOM, "lstc_autotune: counts allocation failed");
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipFree(d_counts);
        return fail(LSTC_ERR_HIP, "lstc_autotune: event creation failed");
    }

    auto measure = [&](const lstc_tune &t, double *ms, std::vector<unsigned long long> &counts) -> lstc_status {
        if (plan) { lstc_batch_plan_destroy(plan); plan = nullptr; }
        LSTC_TRY(lstc_batch_plan_create(graphs, G, &t, &plan));
        LSTC_TRY(lstc_batch_plan_run(plan, (uint64_t *)d_counts, st));  // warm-up
        LSTC_HIP(hipEventRecord(e0, st));
        for (uint32_t r = 0; r < reps; ++r) LSTC_TRY(lstc_batch_plan_run(plan, (uint64_t *)d_counts, st));
        LSTC_HIP(hipEventRecord(e1, st));
        LSTC_HIP(hipEventSynchronize(e1));
        float f = 0;
        LSTC_HIP(hipEventElapsedTime(&f, e0, e1));
        *ms = (double)f / reps;
        LSTC_HIP(hipMemcpy(counts.data(), d_counts, sizeof(unsigned long long) * G, hipMemcpyDeviceToHost));
        return LSTC_OK;
    };

    rc = measure(cur, &cur_ms, ref);
    const uint32_t cand_scale[] = {4, 6, 8, 16};
    const uint32_t cand_small[] = {64, 128, 256};
    const uint32_t cand_group[] = {4, 16, 32, 64};
    const uint32_t cand_grid[] = {4, 6, 8};  // for 4 waves; scaled by 4 / waves below
    const uint32_t cand_unroll[] = {4, 8};
    const uint32_t cand_tiny[] = {1, 4, 8};  // 1 = lane-per-arc path off
    const uint32_t cand_waves[] = {2, 4, 8}; // workgroup size 128 / 256 / 512 (the reference swept 64..512)
    struct Knob { uint32_t lstc_tune::*field; const uint32_t *vals; int nvals; };
    // waves first: grid_per_cu is then searched for the chosen workgroup size
    const Knob knobs[] = {{&lstc_tune::waves, cand_waves, 3},      {&lstc_tune::table_scale, cand_scale, 4},
                          {&lstc_tune::small_max, cand_small, 3},  {&lstc_tune::small_group, cand_group, 4},
                          {&lstc_tune::grid_per_cu, cand_grid, 3}, {&lstc_tune::unroll, cand_unroll, 2},
                          {&lstc_tune::tiny_max, cand_tiny, 3}};
    for (const Knob &k : knobs) {
        if (rc != LSTC_OK) break;
        lstc_tune base = cur;
        for (int i = 0; i < k.nvals && rc == LSTC_OK; ++i) {
            lstc_tune t = base;
            t.*(k.field) = k.vals[i];
            if (k.field == &lstc_tune::grid_per_cu) {  // same waves per CU whatever the workgroup size
                const uint32_t w = t.waves ? t.waves : 4;
                t.grid_per_cu = hmax(1u, k.vals[i] * 4 / w);
            }
            if (memcmp(&t, &cur, sizeof(t)) == 0) continue;
            double ms = 0;
            rc = measure(t, &ms, got);
            if (rc != LSTC_OK) break;
            if (got != ref) {
                rc = fail(LSTC_ERR_INVALID, "lstc_autotune: a candidate changed the counts");
                break;
            }
            if (ms < cur_ms) { cur_ms = ms; cur = t; }
        }
    }
    if (plan) lstc_batch_plan_destroy(plan);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d_counts);
    if (rc != LSTC_OK) return rc;
    // record the values the winning launches ran with, so a cache line is
    // explicit and replays the timed configuration
    LSTC_TRY(lstc_tune_resolve(graphs, G, &cur, best));
    if (best_ms) *best_ms = cur_ms;
    return LSTC_OK;
}

extern "C" lstc_status lstc_tune_cache_lookup(const char *path, uint32_t shape_class, lstc_tune *out) {
    if (!path || !out) return fail(LSTC_ERR_INVALID, "lstc_tune_cache_lookup: NULL argument");
    char dev[128];
    device_key(dev, sizeof(dev));
    FILE *f = fopen(path, "r");
    if (!f) return fail(LSTC_ERR_NOT_READY, "lstc_tune_cache_lookup: no cache at %s", path);
    char line[512];
    bool found = false;
    lstc_tune t{};
    while (fgets(line, sizeof(line), f)) {
        char d[128];
        unsigned cls, a, b, c, g, u, y, w;
        double ms;
        if (sscanf(line, "%127s %u %u %u %u %u %u %u %u %lf", d, &cls, &a, &b, &c, &g, &u, &y, &w, &ms) != 10) continue;
        if (strcmp(d, dev) != 0 || cls != shape_class) continue;
        t = lstc_tune{};
        t.table_scale = a; t.small_max = b; t.small_group = c; t.grid_per_cu = g; t.unroll = u; t.tiny_max = y;
        t.waves = w;
        found = true;  // the last matching line wins
    }
    fclose(f);
    if (!found) return fail(LSTC_ERR_NOT_READY, "lstc_tune_cache_lookup: no entry for class %u", shape_class);
    *out = t;
    return LSTC_OK;
}

extern "C" lstc_status lstc_tune_cache_store(const char *path, uint32_t shape_class, const lstc_tune *t, double ms) {
    if (!path || !t) return fail(LSTC_ERR_INVALID, "lstc_tune_cache_store: NULL argument");
    char dev[128];
    device_key(dev, sizeof(dev));
    FILE *f = fopen(path, "a");
    if (!f) return fail(LSTC_ERR_INVALID, "lstc_tune_cache_store: cannot open %s", path);
    fprintf(f, "%s\t%u\t%u\t%u\t%u\t%u\t%u\t%u\t%u\t%.6f\n", dev, shape_class, t->table_scale, t->small_max,
            t->small_group, t->grid_per_cu, t->unroll, t->tiny_max, t->waves, ms);
    fclose(f);
    return LSTC_OK;
}
