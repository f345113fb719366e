// lstc_cli.h -- shared helpers of the lstc_* command-line tools (host C++).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include <vector>

#include "lstc.h"

#define CLI_CHECK(expr)                                                                        \
    do {                                                                                       \
        lstc_status _s = (expr);                                                               \
        if (_s != LSTC_OK) {                                                                   \
            fprintf(stderr, "%s failed: %s (%s)\n", #expr, lstc_status_string(_s), lstc_last_error()); \
            exit(2);                                                                           \
        }                                                                                      \
    } while (0)

#define CLI_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) {                                                                \
            fprintf(stderr, "%s failed: %s\n", #expr, hipGetErrorString(_e));                  \
            exit(2);                                                                           \
        }                                                                                      \
    } while (0)

namespace lstc_cli {

inline double now_s() {
    timeval t;
    gettimeofday(&t, nullptr);
    return (double)t.tv_sec + 1e-6 * (double)t.tv_usec;
}

template <class T>
inline T *upload(const std::vector<T> &h) {
    T *d = nullptr;
    CLI_HIP(hipMalloc((void **)&d, sizeof(T) * (h.size() ? h.size() : 1)));
    if (!h.empty()) CLI_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return d;
}

template <class T>
inline std::vector<T> download(const T *d, size_t n) {
    std::vector<T> h(n);
    if (n) CLI_HIP(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost));
    return h;
}

struct Workspace {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t need) {
        if (need > bytes) {
            if (p) CLI_HIP(hipFree(p));
            bytes = need < 256 ? 256 : need;
            CLI_HIP(hipMalloc(&p, bytes));
        }
        return p;
    }
    ~Workspace() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace lstc_cli
