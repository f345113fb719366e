// lstc_cli.h -- shared helpers of the lstc_* command-line tools (host C++).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
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

// Owned-source target-side in-lists of an uploaded CSR and the segment words of
// col and the in-lists, built together (lstc_in_lists_segments), so the graph
// is counted split; the caller frees the four arrays.
struct SplitLists {
    uint64_t *in_off = nullptr;
    uint32_t *in_col = nullptr;
    uint64_t *col_seg = nullptr;
    uint64_t *in_seg = nullptr;
    void release() {
        (void)hipFree(in_off);
        (void)hipFree(in_col);
        (void)hipFree(col_seg);
        (void)hipFree(in_seg);
    }
};
inline SplitLists build_in_lists(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint32_t n_owned,
                                 uint64_t m, Workspace &ws) {
    SplitLists sl;
    CLI_HIP(hipMalloc((void **)&sl.in_off, sizeof(uint64_t) * ((size_t)n_rows + 1)));
    CLI_HIP(hipMalloc((void **)&sl.in_col, sizeof(uint32_t) * (m ? m : 1)));
    CLI_HIP(hipMalloc((void **)&sl.col_seg, sizeof(uint64_t) * (m ? m : 1)));
    CLI_HIP(hipMalloc((void **)&sl.in_seg, sizeof(uint64_t) * (m ? m : 1)));
    size_t need = 0;
    CLI_CHECK(lstc_in_lists_workspace(n_rows, n_owned, m, &need));
    void *w = ws.get(need);
    uint64_t entries = 0;
    CLI_CHECK(lstc_in_lists_segments(d_off, d_col, n_rows, n_owned, m, sl.in_off, sl.in_col, sl.col_seg, sl.in_seg,
                                     &entries, w, ws.bytes, nullptr));
    CLI_HIP(hipDeviceSynchronize());
    return sl;
}

// True when argv holds the flag.
inline bool has_flag(int argc, char **argv, int from, const char *flag) {
    for (int i = from; i < argc; ++i)
        if (strcmp(argv[i], flag) == 0) return true;
    return false;
}

}  // namespace lstc_cli
