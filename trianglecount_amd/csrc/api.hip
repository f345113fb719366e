// api.hip -- status strings, per-thread error detail, device properties,
// profiler ranges.
#include <dlfcn.h>
#include <stdarg.h>
#include <stdio.h>
#include <mutex>

#include "lstc_internal.h"

namespace lstc {

static thread_local char g_err[512] = "";

lstc_status fail(lstc_status st, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return st;
}

lstc_status hip_fail(hipError_t e, const char *what) {
    snprintf(g_err, sizeof(g_err), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    return LSTC_ERR_HIP;
}

int device_cus() {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

// Named host ranges around the library's entry points for ROCm profilers
// (rocprofv3 --marker-trace shows them beside the kernels).  The ROCm tracing
// library is loaded at run time, so the library does not link against it;
// without it, or with LSTC_NO_RANGES set, a range costs one branch.
namespace {
struct RangeApi {
    int (*push)(const char *) = nullptr;
    int (*pop)() = nullptr;
    RangeApi() {
        if (getenv("LSTC_NO_RANGES")) return;
        void *h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);  // what rocprofv3 traces
        if (!h) h = dlopen("libroctx64.so.4", RTLD_NOW | RTLD_LOCAL);             // the older tracer's library
        if (!h) return;
        push = (int (*)(const char *))dlsym(h, "roctxRangePushA");
        pop = (int (*)())dlsym(h, "roctxRangePop");
        if (!push || !pop) push = nullptr, pop = nullptr;
    }
};
const RangeApi &range_api() {
    static const RangeApi api;
    return api;
}
}  // namespace

TraceRange::TraceRange(const char *name) : on(range_api().push != nullptr) {
    if (on) range_api().push(name);
}
TraceRange::~TraceRange() {
    if (on) range_api().pop();
}

}  // namespace lstc

extern "C" const char *lstc_status_string(lstc_status st) {
    switch (st) {
        case LSTC_OK: return "ok";
        case LSTC_ERR_INVALID: return "invalid argument";
        case LSTC_ERR_OOM: return "out of device memory";
        case LSTC_ERR_OVER_BUDGET: return "request exceeds the HBM budget";
        case LSTC_ERR_HIP: return "HIP runtime error";
        case LSTC_ERR_WORKSPACE: return "workspace missing or too small";
        case LSTC_ERR_NOT_READY: return "result not ready";
        case LSTC_ERR_CANCELLED: return "request cancelled";
        case LSTC_ERR_DEVICE: return "device-side wait exceeded its bound";
        default: return "unknown status";
    }
}

extern "C" const char *lstc_last_error(void) { return lstc::g_err; }

extern "C" int lstc_version(void) { return 10000; }
