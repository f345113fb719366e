// gen.hip -- seeded synthetic graph generators and the k-way vertex partitions
// (seeded random, contiguous blocks, read from a node-parts file).
// Counter-based (SplitMix64 finaliser) so every sample is a pure function of
// (seed, stream, index); the CPU oracle implements the same definitions.
#include <math.h>
#include <stdio.h>

#include <vector>

#include "lstc_internal.h"

namespace lstc {

static inline uint32_t prob_to_u32(double p) {
    double x = p * 4294967296.0;
    if (x <= 0.0) return 0;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
static inline uint64_t prob_to_u64(double p) {
    if (p <= 0.0) return 0;
    if (p >= 1.0) return UINT64_MAX;
    return (uint64_t)(p * 18446744073709551616.0);
}

// RMAT: one 32-bit uniform per recursion level, two levels per 64-bit draw,
// quadrants a:(0,0) b:(0,1) c:(1,0) d:(1,1), level 0 = most significant bit.  Stream 1.
// Ids are written plus id0 (0 but in the server's grouped preparation, where
// each graph takes its own range of a union's ids).
__global__ void k_gen_rmat(uint32_t scale, uint64_t m_raw, uint32_t tA, uint32_t tAB,
                           uint32_t tABC, uint64_t key, uint32_t id0, uint32_t *__restrict__ src,
                           uint32_t *__restrict__ dst) {
    const uint32_t nh = (scale + 1) / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m_raw;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t s = 0, d = 0;
        uint64_t r = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            if ((l & 1u) == 0) r = rng64(key, i * nh + (l >> 1));
            const uint32_t u = (l & 1u) ? (uint32_t)(r >> 32) : (uint32_t)r;
            const uint32_t bit = 1u << (scale - 1 - l);
            const uint32_t q = (u >= tA) + (u >= tAB) + (u >= tABC);  // 0..3
            if (q & 1u) d |= bit;
            if (q & 2u) s |= bit;
        }
        src[i] = id0 + s;
        dst[i] = id0 + d;
    }
}

// Uniform pairs (G(n, M) sampling).  Stream 2.
__global__ void k_gen_er(uint32_t n, uint64_t m_raw, uint64_t key, uint32_t id0, uint32_t *__restrict__ src,
                         uint32_t *__restrict__ dst) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m_raw;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = rng64(key, i);
        src[i] = id0 + mulhi32((uint32_t)r, n);
        dst[i] = id0 + mulhi32((uint32_t)(r >> 32), n);
    }
}

// Chung-Lu endpoint draws by inverse CDF over an inclusive integer weight prefix.  Stream 3.
__global__ void k_gen_chung_lu(uint32_t n, uint64_t m_raw, uint64_t key,
                               const uint64_t *__restrict__ cw, uint64_t total,
                               uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m_raw;
         e += (uint64_t)gridDim.x * blockDim.x) {
        for (int side = 0; side < 2; ++side) {
            const uint64_t r = rng64(key, 2 * e + (uint64_t)side);
            const uint64_t x = __umul64hi(r, total);
            uint32_t l = 0, h = n;
            while (l < h) {
                const uint32_t mid = l + (h - l) / 2;
                if (cw[mid] > x) h = mid; else l = mid + 1;
            }
            if (side == 0) src[e] = l; else dst[e] = l;
        }
    }
}

// Grid with 3 slots per vertex: right, down, diagonal (present w.p. p via stream 4,
// counter = r*(cols-1)+c).  Missing slots are loops (v, v).
__global__ void k_gen_grid(uint32_t rows, uint32_t cols, uint64_t thr, uint64_t key,
                           uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
    const uint64_t nv = (uint64_t)rows * cols;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(v / cols), c = (uint32_t)(v % cols);
        const uint32_t vv = (uint32_t)v;
        src[3 * v + 0] = vv; dst[3 * v + 0] = (c + 1 < cols) ? vv + 1 : vv;
        src[3 * v + 1] = vv; dst[3 * v + 1] = (r + 1 < rows) ? vv + cols : vv;
        uint32_t dd = vv;
        if (r + 1 < rows && c + 1 < cols) {
            const uint64_t h = rng64(key, (uint64_t)r * (cols - 1) + c);
            if (thr == UINT64_MAX || h < thr) dd = vv + cols + 1;
        }
        src[3 * v + 2] = vv; dst[3 * v + 2] = dd;
    }
}

// part[v] = floor(k * hi32(rng64(stream 5, v)) / 2^32).
__global__ void k_partition(uint32_t n, uint32_t k, uint64_t key, uint32_t *__restrict__ part) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x)
        part[v] = mulhi32((uint32_t)(rng64(key, v) >> 32), k);
}

static uint32_t gen_grid_dim(uint64_t work) {
    return (uint32_t)hmin<uint64_t>((work + 255) / 256, (uint64_t)device_cus() * 32);
}

// Chung-Lu weight table (host): w_i = w_max (1 + i/i0)^(-1/(gamma-1)), i0 solving the
// continuous mean equation by bisection; W_i = floor(w_i * 2^16 + 0.5).
static double cl_mean(double i0, double beta, double n, double w_max) {
    return w_max * i0 / (1.0 - beta) * (pow(1.0 + n / i0, 1.0 - beta) - 1.0) / n;
}
static bool chung_lu_table(uint32_t n, double gamma, double mean_deg, double w_max,
                           std::vector<uint64_t> &cw) {
    if (n == 0 || !(gamma > 1.0) || fabs(gamma - 2.0) < 1e-9) return false;
    const double beta = 1.0 / (gamma - 1.0);
    double lo = 1e-9, hi = (double)n;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (cl_mean(mid, beta, (double)n, w_max) < mean_deg) lo = mid; else hi = mid;
    }
    const double i0 = 0.5 * (lo + hi);
    cw.resize(n);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double w = w_max * pow(1.0 + (double)i / i0, -beta);
        acc += (uint64_t)(w * 65536.0 + 0.5);
        cw[i] = acc;
    }
    return true;
}

lstc_status gen_rmat_launch(uint32_t scale, uint64_t m_raw, double a, double b, double c, uint64_t seed, uint32_t id0,
                            uint32_t *d_src, uint32_t *d_dst, hipStream_t s) {
    if (scale == 0 || scale > 31) return fail(LSTC_ERR_INVALID, "lstc_gen_rmat: scale %u not in 1..31", scale);
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
        return fail(LSTC_ERR_INVALID, "lstc_gen_rmat: bad a/b/c");
    if ((uint64_t)id0 + (1ull << scale) > 0xFFFFFFFFull) return fail(LSTC_ERR_INVALID, "lstc_gen_rmat: ids overflow");
    if (m_raw == 0) return LSTC_OK;
    if (!d_src || !d_dst) return fail(LSTC_ERR_INVALID, "lstc_gen_rmat: NULL output");
    hipLaunchKernelGGL(k_gen_rmat, dim3(gen_grid_dim(m_raw)), dim3(256), 0, s, scale, m_raw, prob_to_u32(a),
                       prob_to_u32(a + b), prob_to_u32(a + b + c), stream_key(seed, 1), id0, d_src, d_dst);
    LSTC_LAUNCHED("k_gen_rmat");
    return LSTC_OK;
}

lstc_status gen_er_launch(uint32_t n, uint64_t m_raw, uint64_t seed, uint32_t id0, uint32_t *d_src, uint32_t *d_dst,
                          hipStream_t s) {
    if (n == 0 || n == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_gen_er: bad n");
    if ((uint64_t)id0 + n > 0xFFFFFFFFull) return fail(LSTC_ERR_INVALID, "lstc_gen_er: ids overflow");
    if (m_raw == 0) return LSTC_OK;
    if (!d_src || !d_dst) return fail(LSTC_ERR_INVALID, "lstc_gen_er: NULL output");
    hipLaunchKernelGGL(k_gen_er, dim3(gen_grid_dim(m_raw)), dim3(256), 0, s, n, m_raw, stream_key(seed, 2), id0,
                       d_src, d_dst);
    LSTC_LAUNCHED("k_gen_er");
    return LSTC_OK;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_gen_rmat(uint32_t scale, uint64_t m_raw, double a, double b, double c,
                                     uint64_t seed, uint32_t *d_src, uint32_t *d_dst,
                                     lstc_stream s) {
    TraceRange range_("lstc_gen_rmat");
    return gen_rmat_launch(scale, m_raw, a, b, c, seed, 0, d_src, d_dst, (hipStream_t)s);
}

extern "C" lstc_status lstc_gen_er(uint32_t n, uint64_t m_raw, uint64_t seed, uint32_t *d_src,
                                   uint32_t *d_dst, lstc_stream s) {
    TraceRange range_("lstc_gen_er");
    return gen_er_launch(n, m_raw, seed, 0, d_src, d_dst, (hipStream_t)s);
}

extern "C" lstc_status lstc_gen_chung_lu_workspace(uint32_t n, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_gen_chung_lu_workspace: NULL bytes");
    *bytes = ((size_t)n * sizeof(uint64_t) + 255) & ~size_t(255);
    return LSTC_OK;
}

extern "C" lstc_status lstc_gen_chung_lu(uint32_t n, double gamma, double mean_deg, double w_max,
                                         uint64_t m_raw, uint64_t seed, uint32_t *d_src,
                                         uint32_t *d_dst, void *d_ws, size_t ws_bytes,
                                         lstc_stream s) {
    TraceRange range_("lstc_gen_chung_lu");
    std::vector<uint64_t> cw;
    if (n == 0xFFFFFFFFu || !chung_lu_table(n, gamma, mean_deg, w_max, cw))
        return fail(LSTC_ERR_INVALID, "lstc_gen_chung_lu: bad parameters");
    if ((size_t)n * 8 > ws_bytes || !d_ws)
        return fail(LSTC_ERR_WORKSPACE, "lstc_gen_chung_lu: workspace too small");
    if (m_raw == 0) return LSTC_OK;
    if (!d_src || !d_dst) return fail(LSTC_ERR_INVALID, "lstc_gen_chung_lu: NULL output");
    hipStream_t st = (hipStream_t)s;
    uint64_t *d_cw = (uint64_t *)d_ws;
    LSTC_HIP(hipMemcpyAsync(d_cw, cw.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gen_chung_lu, dim3(gen_grid_dim(m_raw)), dim3(256), 0, st, n, m_raw,
                       stream_key(seed, 3), d_cw, cw.back(), d_src, d_dst);
    LSTC_LAUNCHED("k_gen_chung_lu");
    // the host table must outlive the copy
    LSTC_HIP(hipStreamSynchronize(st));
    return LSTC_OK;
}

extern "C" lstc_status lstc_gen_grid(uint32_t rows, uint32_t cols, double p_diag, uint64_t seed,
                                     uint32_t *d_src, uint32_t *d_dst, lstc_stream s) {
    TraceRange range_("lstc_gen_grid");
    const uint64_t nv = (uint64_t)rows * cols;
    if (nv >= 0xFFFFFFFFull) return fail(LSTC_ERR_INVALID, "lstc_gen_grid: too many vertices");
    if (nv == 0) return LSTC_OK;
    if (!d_src || !d_dst) return fail(LSTC_ERR_INVALID, "lstc_gen_grid: NULL output");
    hipLaunchKernelGGL(k_gen_grid, dim3(gen_grid_dim(nv)), dim3(256), 0, (hipStream_t)s, rows, cols,
                       prob_to_u64(p_diag), stream_key(seed, 4), d_src, d_dst);
    LSTC_LAUNCHED("k_gen_grid");
    return LSTC_OK;
}

// Contiguous blocks (preprocess_GPU/generate_clean_undir_dataset/
// Halo_creation_preprocess.py:292-300): l = n // k, part i holds [i*l, (i+1)*l),
// the remainder (and everything when n < k) stays in part 0.
__global__ void k_partition_contig(uint32_t n, uint32_t k, uint32_t l, uint32_t *__restrict__ part) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = l ? v / l : 0;
        part[v] = i < k ? (uint32_t)i : 0u;
    }
}

extern "C" lstc_status lstc_partition_contig(uint32_t n, uint32_t k, uint32_t *d_part, lstc_stream s) {
    if (k == 0) return fail(LSTC_ERR_INVALID, "lstc_partition_contig: k must be > 0");
    if (n == 0) return LSTC_OK;
    if (!d_part) return fail(LSTC_ERR_INVALID, "lstc_partition_contig: NULL output");
    hipLaunchKernelGGL(k_partition_contig, dim3(gen_grid_dim(n)), dim3(256), 0, (hipStream_t)s, n, k, n / k, d_part);
    LSTC_LAUNCHED("k_partition_contig");
    return LSTC_OK;
}

// A node-parts file (one integer per vertex, whitespace-separated: the format
// Halo_creation_preprocess.py:308-312 writes and round_clean_subgraph.cu:964-968
// reads), e.g. an assignment computed by METIS outside this library.  The first
// n values are read (as the reference reads them); a value >= k is refused,
// since that vertex would belong to no part.
extern "C" lstc_status lstc_partition_read(const char *path, uint32_t n, uint32_t k, uint32_t *d_part,
                                           lstc_stream s) {
    if (!path) return fail(LSTC_ERR_INVALID, "lstc_partition_read: NULL path");
    if (k == 0) return fail(LSTC_ERR_INVALID, "lstc_partition_read: k must be > 0");
    if (n && !d_part) return fail(LSTC_ERR_INVALID, "lstc_partition_read: NULL output");
    FILE *f = fopen(path, "r");
    if (!f) return fail(LSTC_ERR_INVALID, "lstc_partition_read: cannot open %s", path);
    std::vector<uint32_t> part(n);
    for (uint32_t v = 0; v < n; ++v) {
        unsigned long long x = 0;
        if (fscanf(f, "%llu", &x) != 1) {
            fclose(f);
            return fail(LSTC_ERR_INVALID, "lstc_partition_read: %s holds %u values, %u vertices expected", path, v, n);
        }
        if (x >= k) {
            fclose(f);
            return fail(LSTC_ERR_INVALID, "lstc_partition_read: vertex %u has part %llu >= k = %u", v, x, k);
        }
        part[v] = (uint32_t)x;
    }
    fclose(f);
    if (n == 0) return LSTC_OK;
    LSTC_HIP(hipMemcpyAsync(d_part, part.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, (hipStream_t)s));
    LSTC_HIP(hipStreamSynchronize((hipStream_t)s));  // the host buffer goes out of scope
    return LSTC_OK;
}

extern "C" lstc_status lstc_partition_random(uint32_t n, uint32_t k, uint64_t seed,
                                             uint32_t *d_part, lstc_stream s) {
    if (k == 0) return fail(LSTC_ERR_INVALID, "lstc_partition_random: k must be > 0");
    if (n == 0) return LSTC_OK;
    if (!d_part) return fail(LSTC_ERR_INVALID, "lstc_partition_random: NULL output");
    hipLaunchKernelGGL(k_partition, dim3(gen_grid_dim(n)), dim3(256), 0, (hipStream_t)s, n, k,
                       stream_key(seed, 5), d_part);
    LSTC_LAUNCHED("k_partition");
    return LSTC_OK;
}
