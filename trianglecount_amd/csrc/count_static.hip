// count_static.hip -- lstc_count_split_static: the split count with no
// allocation and no workspace argument.
//
// For callers that cannot allocate device memory at the call (code under
// stream capture, tight serving loops, harnesses that forbid it): the engine's
// scratch (plan control words, per-wave plan counts, work queue) lives in a
// file-scope device array of kStaticScratch bytes, one copy per GPU, held until
// the process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept; host code passes
// that address, never the array's name, to the kernels.  The plan kernels
// rewrite every scratch word a count reads before reading it (control words
// and counts are zeroed by the plan's first kernel), so nothing has to be
// cleared between calls.  Calls that overlap on one GPU share the array and
// race: run them on one stream, or serialise them.
#include <mutex>

#include "count_engine.h"

namespace lstc {

static constexpr size_t kStaticScratch = size_t(32) << 20;  // 32 MiB per GPU
static constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratch];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// The device address of this device's scratch (looked up on the first call on each device).
static lstc_status scratch_ptr(void **out) {
    int dev = 0;
    LSTC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(LSTC_ERR_INVALID, "lstc_count_split_static: device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        LSTC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(LSTC_ERR_HIP, "lstc_count_split_static: no scratch address on device %d", dev);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return LSTC_OK;
}

}  // namespace lstc

using namespace lstc;

extern "C" size_t lstc_count_split_static_limit(void) { return kStaticScratch; }

extern "C" lstc_status lstc_count_split_static(const uint64_t *d_off, const uint32_t *d_col, const uint64_t *d_in_off,
                                               const uint32_t *d_in_col, const uint64_t *d_col_seg,
                                               const uint64_t *d_in_seg, uint32_t n_rows, uint32_t n_owned,
                                               uint64_t m, uint64_t *d_count, const lstc_tune *tune, lstc_stream s) {
    TraceRange range_("lstc_count_split_static");
    if (!d_count) return fail(LSTC_ERR_INVALID, "lstc_count_split_static: NULL output");
    if (!d_in_off || !d_in_col) return fail(LSTC_ERR_INVALID, "lstc_count_split_static: NULL in-lists");
    HostGraph h{d_off, d_col, nullptr, n_rows, n_owned, m, d_in_off, d_in_col, d_col_seg, d_in_seg};
    LSTC_TRY(validate_graph(h, "lstc_count_split_static"));
    const size_t need = count_ws_bytes(1, n_rows, m);
    if (need > kStaticScratch)
        return fail(LSTC_ERR_INVALID,
                    "lstc_count_split_static: this graph needs %zu bytes of scratch, the static scratch holds %zu "
                    "(use lstc_count_split with a workspace)", need, kStaticScratch);
    void *ws = nullptr;
    LSTC_TRY(scratch_ptr(&ws));
    return count_launch(&h, 1, nullptr, nullptr, (unsigned long long *)d_count, tune, ws, kStaticScratch,
                        (hipStream_t)s, /*timed=*/false);
}
