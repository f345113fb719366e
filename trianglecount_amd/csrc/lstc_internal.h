// lstc_internal.h -- shared internals of liblstc (gfx950 / wave64 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "lstc.h"

namespace lstc {

// ---------------------------------------------------------------- errors
// Records a detail message for lstc_last_error() (thread-local) and returns st.
lstc_status fail(lstc_status st, const char *fmt, ...);
lstc_status hip_fail(hipError_t e, const char *what);

// A named host range for ROCm profilers for the lifetime of the object (api.hip).
struct TraceRange {
    explicit TraceRange(const char *name);
    ~TraceRange();
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
    bool on;
};

#define LSTC_HIP(expr)                                                  \
    do {                                                                \
        hipError_t _e = (expr);                                         \
        if (_e != hipSuccess) return ::lstc::hip_fail(_e, #expr);       \
    } while (0)

#define LSTC_TRY(expr)                                                  \
    do {                                                                \
        lstc_status _s = (expr);                                        \
        if (_s != LSTC_OK) return _s;                                   \
    } while (0)

// Launch check: kernel launch errors surface through hipGetLastError.
#define LSTC_LAUNCHED(name)                                             \
    do {                                                                \
        hipError_t _e = hipGetLastError();                              \
        if (_e != hipSuccess) return ::lstc::hip_fail(_e, name);        \
    } while (0)

// ---------------------------------------------------------------- workspace carving
struct Carver {
    char *base;
    size_t cap, used;
    explicit Carver(void *b, size_t c) : base((char *)b), cap(c), used(0) {}
    template <class T> T *take(size_t count) {
        size_t off = (used + 255) & ~size_t(255);
        used = off + count * sizeof(T);
        return (base && used <= cap) ? (T *)(base + off) : nullptr;
    }
    bool ok() const { return used <= cap; }
    size_t need() const { return (used + 255) & ~size_t(255); }
};

// ---------------------------------------------------------------- RNG (SplitMix64 finaliser)
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(mix64(seed + 0x9E3779B97F4A7C15ull) ^ (stream * 0xD1B54A32D192ED03ull));
}
__host__ __device__ inline uint64_t rng64(uint64_t key, uint64_t ctr) {
    return mix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ull);
}
__host__ __device__ inline uint32_t mulhi32(uint32_t r, uint32_t n) {
    return (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

// ---------------------------------------------------------------- launch helpers
constexpr int kWave = 64;

inline uint32_t ceil_div_u32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
template <class T> inline T hmin(T a, T b) { return a < b ? a : b; }
template <class T> inline T hmax(T a, T b) { return a < b ? b : a; }

// Number of compute units of the current device (cached per device).
int device_cus();

// ---------------------------------------------------------------- device primitives (prims.hip)
// Exclusive scan of n uint64 values in place-compatible form (in may equal out).
// Returns the total through d_total (device, may be null).  Workspace from scan_ws_bytes.
size_t scan_ws_bytes(uint64_t n);
lstc_status scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *d_total,
                               void *ws, size_t ws_bytes, hipStream_t s);
// Same over uint32 input, uint64 output.
lstc_status scan_exclusive_u32_to_u64(const uint32_t *in, uint64_t *out, uint64_t n,
                                      uint64_t *d_total, void *ws, size_t ws_bytes, hipStream_t s);

// LSD radix sort of uint64 keys over the low key_bits bits.  keys and alt must hold n
// entries; the result ends in *result (either keys or alt).
size_t radix_sort_ws_bytes(uint64_t n);
lstc_status radix_sort_u64(uint64_t *keys, uint64_t *alt, uint64_t n, int key_bits,
                           uint64_t **result, void *ws, size_t ws_bytes, hipStream_t s,
                           bool first_counts_ready = false);
// The sort's passes work on blocks of kRadixBlock consecutive keys.  A kernel
// that writes the keys block by block can count the first digit (the low 8
// bits) itself: the count of digit d in block t goes to
// radix_first_counts(ws, n)[d * radix_blocks(n) + t], and the sort is then
// called with first_counts_ready, which skips its first counting pass.
constexpr uint64_t kRadixBlock = 16384;
inline uint64_t radix_blocks(uint64_t n) { return (n + kRadixBlock - 1) / kRadixBlock; }
uint32_t *radix_first_counts(void *ws, uint64_t n);

// (a) clean + orient without synchronising (prep.hip); sizes stay on the device:
// (*d_sizes)[1] = vertices after cleaning, (*d_sizes)[2] = undirected edges,
// (*d_sizes)[3] = in-list entries, (*d_sizes)[4] = streaming copy entries.
// In-lists and segment words are optional (NULL); segment words of the in-lists
// need the in-lists.  d_scol (optional, with segment words): a streaming copy of
// the columns whose rows start every 32 entries; the segment words then hold
// row starts in it.
lstc_status clean_orient_launch(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw, uint32_t n,
                                int compact, uint64_t *d_sym_off, uint32_t *d_sym_col, uint64_t *d_off,
                                uint32_t *d_col, uint64_t *d_in_off, uint32_t *d_in_col, uint64_t *d_col_seg,
                                uint64_t *d_in_seg, void *d_ws, size_t ws_bytes, hipStream_t st,
                                const uint64_t **d_sizes, const uint32_t **d_bad, uint32_t *d_scol = nullptr);

// Seeded generators (gen.hip) writing every id plus id0: lstc_gen_rmat and
// lstc_gen_er are these with id0 = 0; the server's grouped preparation gives
// each graph its own id range of one union.
lstc_status gen_rmat_launch(uint32_t scale, uint64_t m_raw, double a, double b, double c, uint64_t seed, uint32_t id0,
                            uint32_t *d_src, uint32_t *d_dst, hipStream_t s);
lstc_status gen_er_launch(uint32_t n, uint64_t m_raw, uint64_t seed, uint32_t id0, uint32_t *d_src, uint32_t *d_dst,
                          hipStream_t s);

// Fill kernels (used instead of hipMemsetAsync so callers may capture them in graphs).
lstc_status fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);
lstc_status fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
lstc_status fill_u8(uint8_t *p, uint8_t v, uint64_t n, hipStream_t s);

}  // namespace lstc
