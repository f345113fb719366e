// prep.hip -- (a) GPU clean + orient into a uint32 CSR with uint64 offsets.
//
// Replaces the DGL cleaning of Med_dataset/without_part_preprocess.py:193-208
// (remove_self_loop, to_bidirected, isolated-node removal) and the CuPy
// orientation Find_size / cp.cumsum / Convert of :19-111,253-270.
//
// Pipeline (all stream-ordered, sizes kept on the device until the end):
//   keys  : each raw pair (s, d), s != d, emits (s<<b | d) and (d<<b | s);
//           loops and out-of-range ids emit a sentinel that sorts last
//   sort  : LSD radix sort over 2b bits (prims.hip)
//   unique: flag first-of-run, scan, scatter  -> symmetric arc list (sorted)
//   degree: per-source atomic count           -> deg[v]
//   relabel (compact): id' = rank of v among vertices with deg > 0 (monotone)
//   symmetric CSR: offsets = scan(deg) compacted, columns relabelled
//   orient: keep s->d iff deg s < deg d or (equal and d < s); flag, scan,
//           scatter keeps each row ascending; offsets = scan of kept per row.
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint64_t kSentinel = ~0ull;

static int bits_for(uint64_t n) {  // bits to represent values < n
    int b = 0;
    while (b < 64 && ((n - 1) >> b)) ++b;
    return b == 0 ? 1 : b;
}

#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

__global__ void k_make_keys(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                            uint64_t m_raw, uint32_t n, uint32_t b, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ bad) {
    GRID_STRIDE(i, m_raw) {
        const uint32_t s = src[i], d = dst[i];
        uint64_t k0 = kSentinel, k1 = kSentinel;
        if (s >= n || d >= n) {
            *bad = 1;
        } else if (s != d) {
            k0 = ((uint64_t)s << b) | d;
            k1 = ((uint64_t)d << b) | s;
        }
        keys[2 * i] = k0;
        keys[2 * i + 1] = k1;
    }
}

__global__ void k_unique_flags(const uint64_t *__restrict__ k, uint64_t nk, uint32_t *__restrict__ f) {
    GRID_STRIDE(i, nk) {
        const uint64_t x = k[i];
        f[i] = (x != kSentinel && (i == 0 || k[i - 1] != x)) ? 1u : 0u;
    }
}

__global__ void k_unique_scatter(const uint64_t *__restrict__ k, uint64_t nk,
                                 const uint32_t *__restrict__ f, const uint64_t *__restrict__ pos,
                                 uint64_t *__restrict__ uk) {
    GRID_STRIDE(i, nk) if (f[i]) uk[pos[i]] = k[i];
}

// Row boundaries of the sorted symmetric arc list: first[s], last[s] (exclusive).
// Vertices without arcs keep first == last == 0 (pre-filled).  No atomics: a hub
// row is one run of the sorted list, so its degree is last - first.
__global__ void k_row_bounds(const uint64_t *__restrict__ uk, const uint64_t *__restrict__ d_m2, uint32_t b,
                             uint64_t *__restrict__ first, uint64_t *__restrict__ last) {
    const uint64_t m2 = *d_m2;
    GRID_STRIDE(i, m2) {
        const uint64_t s = uk[i] >> b;
        if (i == 0 || (uk[i - 1] >> b) != s) first[s] = i;
        if (i + 1 == m2 || (uk[i + 1] >> b) != s) last[s] = i + 1;
    }
}

__global__ void k_degree(const uint64_t *__restrict__ first, const uint64_t *__restrict__ last, uint32_t n,
                         uint32_t *__restrict__ deg) {
    GRID_STRIDE(v, n) deg[v] = (uint32_t)(last[v] - first[v]);
}

// kept arcs per row from the exclusive scan of keep flags at the row boundaries
__global__ void k_out_degree(const uint64_t *__restrict__ first, const uint64_t *__restrict__ last,
                             const uint64_t *__restrict__ pos, uint32_t n, uint32_t *__restrict__ dout) {
    GRID_STRIDE(v, n) dout[v] = (uint32_t)(pos[last[v]] - pos[first[v]]);
}

__global__ void k_nonzero(const uint32_t *__restrict__ deg, uint32_t n, int compact,
                          uint32_t *__restrict__ f) {
    GRID_STRIDE(v, n) f[v] = (!compact || deg[v] != 0) ? 1u : 0u;
}

// nid[v] = new id (from the exclusive scan of the keep flags); writes compacted offsets.
__global__ void k_relabel_offsets(const uint32_t *__restrict__ keepv, const uint64_t *__restrict__ nidscan,
                                  const uint64_t *__restrict__ degscan, uint32_t n,
                                  const uint64_t *__restrict__ d_n_out, const uint64_t *__restrict__ d_m2,
                                  uint32_t *__restrict__ nid, uint64_t *__restrict__ sym_off) {
    GRID_STRIDE(v, n) {
        if (keepv[v]) {
            const uint32_t id = (uint32_t)nidscan[v];
            nid[v] = id;
            if (sym_off) sym_off[id] = degscan[v];
        } else {
            nid[v] = 0xFFFFFFFFu;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && sym_off) sym_off[*d_n_out] = *d_m2;
}

__global__ void k_sym_cols(const uint64_t *__restrict__ uk, const uint64_t *__restrict__ d_m2,
                           uint64_t mask, const uint32_t *__restrict__ nid, uint32_t *__restrict__ sym_col) {
    const uint64_t m2 = *d_m2;
    GRID_STRIDE(i, m2) sym_col[i] = nid[uk[i] & mask];
}

// keep flags for [0, nk]; entries from m2 on (stale from the unique pass) become 0
__global__ void k_orient_flags(const uint64_t *__restrict__ uk, const uint64_t *__restrict__ d_m2,
                               uint64_t nk, uint32_t b, uint64_t mask, const uint32_t *__restrict__ deg,
                               uint32_t *__restrict__ f) {
    const uint64_t m2 = *d_m2;
    GRID_STRIDE(i, nk + 1) {
        if (i >= m2) { f[i] = 0u; continue; }
        const uint64_t k = uk[i];
        const uint32_t s = (uint32_t)(k >> b), d = (uint32_t)(k & mask);
        const uint32_t ds = deg[s], dd = deg[d];
        f[i] = (ds < dd || (ds == dd && d < s)) ? 1u : 0u;
    }
}

// In-arcs d -> s (the reverse keys of row s) with d+(d) <= d+(s): the arcs the
// count handles from the target side.  Flags for [0, nk]; zero from m2 on.
__global__ void k_in_flags(const uint64_t *__restrict__ uk, const uint64_t *__restrict__ d_m2, uint64_t nk,
                           uint32_t b, uint64_t mask, const uint32_t *__restrict__ deg,
                           const uint32_t *__restrict__ dout, uint32_t *__restrict__ f) {
    const uint64_t m2 = *d_m2;
    GRID_STRIDE(i, nk + 1) {
        if (i >= m2) { f[i] = 0u; continue; }
        const uint64_t k = uk[i];
        const uint32_t s = (uint32_t)(k >> b), d = (uint32_t)(k & mask);
        const uint32_t ds = deg[s], dd = deg[d];
        const bool out_arc = ds < dd || (ds == dd && d < s);
        uint32_t fl = 0;
        if (!out_arc) fl = dout[d] <= dout[s] ? 1u : 0u;  // out-degrees only read for in-arcs
        f[i] = fl;
    }
}

__global__ void k_orient_scatter(const uint64_t *__restrict__ uk, const uint64_t *__restrict__ d_m2,
                                 uint64_t mask, const uint32_t *__restrict__ f,
                                 const uint64_t *__restrict__ pos, const uint32_t *__restrict__ nid,
                                 uint32_t *__restrict__ col) {
    const uint64_t m2 = *d_m2;
    GRID_STRIDE(i, m2) if (f[i]) col[pos[i]] = nid[uk[i] & mask];
}

__global__ void k_orient_offsets(const uint32_t *__restrict__ keepv, const uint64_t *__restrict__ nidscan,
                                 const uint64_t *__restrict__ doscan, uint32_t n,
                                 const uint64_t *__restrict__ d_n_out, const uint64_t *__restrict__ d_m,
                                 uint64_t *__restrict__ off) {
    GRID_STRIDE(v, n) if (keepv[v]) off[nidscan[v]] = doscan[v];
    if (blockIdx.x == 0 && threadIdx.x == 0) off[*d_n_out] = *d_m;
}

struct PrepLayout {
    uint64_t *keys, *alt, *pos;
    uint32_t *flags;
    uint32_t *deg, *keepv, *nid, *dout, *din;
    uint64_t *degscan, *nidscan, *doscan, *first, *last, *discan;
    uint64_t *scal;  // [0]=m2 [1]=n_out [2]=m [3]=unused
    uint32_t *bad;
    void *sort_ws;
    size_t sort_bytes;
    void *scan_ws;
    size_t scan_bytes;
    size_t bytes;
};

static PrepLayout prep_layout(void *ws, size_t ws_bytes, uint32_t n, uint64_t m_raw) {
    Carver c(ws, ws_bytes);
    PrepLayout L;
    const uint64_t nk = 2 * m_raw;
    L.keys = c.take<uint64_t>(nk + 1);
    L.alt = c.take<uint64_t>(nk + 1);
    L.pos = c.take<uint64_t>(nk + 1);
    L.flags = c.take<uint32_t>(nk + 1);
    L.deg = c.take<uint32_t>((size_t)n + 1);
    L.keepv = c.take<uint32_t>((size_t)n + 1);
    L.nid = c.take<uint32_t>((size_t)n + 1);
    L.dout = c.take<uint32_t>((size_t)n + 1);
    L.din = c.take<uint32_t>((size_t)n + 1);
    L.discan = c.take<uint64_t>((size_t)n + 1);
    L.degscan = c.take<uint64_t>((size_t)n + 1);
    L.nidscan = c.take<uint64_t>((size_t)n + 1);
    L.doscan = c.take<uint64_t>((size_t)n + 1);
    L.first = c.take<uint64_t>((size_t)n + 1);
    L.last = c.take<uint64_t>((size_t)n + 1);
    L.scal = c.take<uint64_t>(8);
    L.bad = c.take<uint32_t>(4);
    L.sort_bytes = radix_sort_ws_bytes(nk);
    L.sort_ws = c.take<char>(L.sort_bytes);
    L.scan_bytes = scan_ws_bytes(nk + 1 > n ? nk + 1 : n);
    L.scan_ws = c.take<char>(L.scan_bytes);
    L.bytes = c.need();
    return L;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_clean_orient_workspace(uint32_t n, uint64_t m_raw, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_clean_orient_workspace: NULL bytes");
    *bytes = prep_layout(nullptr, 0, n, m_raw).bytes;
    return LSTC_OK;
}

// Enqueues the whole pipeline without synchronising.  On return *d_sizes points
// at device words {m2, n_out, m, -} and *d_bad at the out-of-range flag; both are
// valid once the stream reaches this point.
lstc_status lstc::clean_orient_launch(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw, uint32_t n,
                                      int compact, uint64_t *d_sym_off, uint32_t *d_sym_col, uint64_t *d_off,
                                      uint32_t *d_col, uint64_t *d_in_off, uint32_t *d_in_col, void *d_ws,
                                      size_t ws_bytes, hipStream_t st,
                                      const uint64_t **d_sizes, const uint32_t **d_bad) {
    if (n == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: n must be < 2^32-1");
    if (!d_off) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL output");
    if ((d_sym_off == nullptr) != (d_sym_col == nullptr))
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient: give both symmetric outputs or neither");
    if (m_raw && (!d_src || !d_dst || !d_col)) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL edge arrays");
    PrepLayout L = prep_layout(d_ws, ws_bytes, n, m_raw);
    if (!d_ws || ws_bytes < L.bytes)
        return fail(LSTC_ERR_WORKSPACE, "lstc_clean_orient: workspace %zu < %zu bytes", ws_bytes, L.bytes);
    const uint32_t b = (uint32_t)bits_for(n ? n : 1);
    const uint64_t mask = (b >= 64) ? ~0ull : ((1ull << b) - 1);
    const uint64_t nk = 2 * m_raw;
    const uint32_t grid_k = (uint32_t)hmin<uint64_t>((nk + 255) / 256 + 1, (uint64_t)device_cus() * 32);
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n + 255) / 256 + 1, (uint64_t)device_cus() * 32);

    LSTC_TRY(fill_u64(L.scal, 0, 8, st));
    LSTC_TRY(fill_u32(L.bad, 0, 4, st));
    LSTC_TRY(fill_u32(L.deg, 0, (uint64_t)n + 1, st));
    LSTC_TRY(fill_u32(L.dout, 0, (uint64_t)n + 1, st));
    if (m_raw) {
        hipLaunchKernelGGL(k_make_keys, dim3(grid_k), dim3(256), 0, st, d_src, d_dst, m_raw, n, b, L.keys, L.bad);
        LSTC_LAUNCHED("k_make_keys");
        uint64_t *sorted = nullptr;
        LSTC_TRY(radix_sort_u64(L.keys, L.alt, nk, (int)(2 * b > 64 ? 64 : 2 * b), &sorted, L.sort_ws, L.sort_bytes, st));
        uint64_t *uk = (sorted == L.keys) ? L.alt : L.keys;
        hipLaunchKernelGGL(k_unique_flags, dim3(grid_k), dim3(256), 0, st, sorted, nk, L.flags);
        LSTC_LAUNCHED("k_unique_flags");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.flags, L.pos, nk, L.scal + 0, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_unique_scatter, dim3(grid_k), dim3(256), 0, st, sorted, nk, L.flags, L.pos, uk);
        LSTC_LAUNCHED("k_unique_scatter");
        LSTC_TRY(fill_u64(L.first, 0, (uint64_t)n + 1, st));
        LSTC_TRY(fill_u64(L.last, 0, (uint64_t)n + 1, st));
        hipLaunchKernelGGL(k_row_bounds, dim3(grid_k), dim3(256), 0, st, uk, L.scal + 0, b, L.first, L.last);
        LSTC_LAUNCHED("k_row_bounds");
        hipLaunchKernelGGL(k_degree, dim3(grid_n), dim3(256), 0, st, L.first, L.last, n, L.deg);
        LSTC_LAUNCHED("k_degree");
        hipLaunchKernelGGL(k_nonzero, dim3(grid_n), dim3(256), 0, st, L.deg, n, compact, L.keepv);
        LSTC_LAUNCHED("k_nonzero");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.keepv, L.nidscan, n, L.scal + 1, L.scan_ws, L.scan_bytes, st));
        LSTC_TRY(scan_exclusive_u32_to_u64(L.deg, L.degscan, n, nullptr, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_relabel_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.degscan, n,
                           L.scal + 1, L.scal + 0, L.nid, d_sym_off);
        LSTC_LAUNCHED("k_relabel_offsets");
        if (d_sym_col) {
            hipLaunchKernelGGL(k_sym_cols, dim3(grid_k), dim3(256), 0, st, uk, L.scal + 0, mask, L.nid, d_sym_col);
            LSTC_LAUNCHED("k_sym_cols");
        }
        // k_orient_flags also zeroes the flags in [m2, nk] left over from the unique pass
        hipLaunchKernelGGL(k_orient_flags, dim3(grid_k), dim3(256), 0, st, uk, L.scal + 0, nk, b, mask, L.deg,
                           L.flags);
        LSTC_LAUNCHED("k_orient_flags");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.flags, L.pos, nk + 1, L.scal + 2, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_out_degree, dim3(grid_n), dim3(256), 0, st, L.first, L.last, L.pos, n, L.dout);
        LSTC_LAUNCHED("k_out_degree");
        hipLaunchKernelGGL(k_orient_scatter, dim3(grid_k), dim3(256), 0, st, uk, L.scal + 0, mask, L.flags,
                           L.pos, L.nid, d_col);
        LSTC_LAUNCHED("k_orient_scatter");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.dout, L.doscan, n, nullptr, L.scan_ws, L.scan_bytes, st));
        hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                           L.scal + 1, L.scal + 2, d_off);
        LSTC_LAUNCHED("k_orient_offsets");
        if (d_in_off) {
            // in-lists of the arcs counted from the target side (same sorted key list, no atomics)
            hipLaunchKernelGGL(k_in_flags, dim3(grid_k), dim3(256), 0, st, uk, L.scal + 0, nk, b, mask, L.deg,
                               L.dout, L.flags);
            LSTC_LAUNCHED("k_in_flags");
            LSTC_TRY(scan_exclusive_u32_to_u64(L.flags, L.pos, nk + 1, L.scal + 3, L.scan_ws, L.scan_bytes, st));
            hipLaunchKernelGGL(k_orient_scatter, dim3(grid_k), dim3(256), 0, st, uk, L.scal + 0, mask, L.flags,
                               L.pos, L.nid, d_in_col);
            LSTC_LAUNCHED("k_orient_scatter(in)");
            hipLaunchKernelGGL(k_out_degree, dim3(grid_n), dim3(256), 0, st, L.first, L.last, L.pos, n, L.din);
            LSTC_LAUNCHED("k_out_degree(in)");
            LSTC_TRY(scan_exclusive_u32_to_u64(L.din, L.discan, n, nullptr, L.scan_ws, L.scan_bytes, st));
            hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.discan, n,
                               L.scal + 1, L.scal + 3, d_in_off);
            LSTC_LAUNCHED("k_orient_offsets(in)");
        }
    } else {
        // no edges: every vertex is isolated
        hipLaunchKernelGGL(k_nonzero, dim3(grid_n), dim3(256), 0, st, L.deg, n, compact, L.keepv);
        LSTC_LAUNCHED("k_nonzero");
        LSTC_TRY(scan_exclusive_u32_to_u64(L.keepv, L.nidscan, n, L.scal + 1, L.scan_ws, L.scan_bytes, st));
        LSTC_TRY(fill_u64(L.doscan, 0, (uint64_t)n + 1, st));
        hipLaunchKernelGGL(k_relabel_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                           L.scal + 1, L.scal + 0, L.nid, d_sym_off);
        LSTC_LAUNCHED("k_relabel_offsets");
        hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                           L.scal + 1, L.scal + 2, d_off);
        LSTC_LAUNCHED("k_orient_offsets");
        if (d_in_off) {
            hipLaunchKernelGGL(k_orient_offsets, dim3(grid_n), dim3(256), 0, st, L.keepv, L.nidscan, L.doscan, n,
                               L.scal + 1, L.scal + 3, d_in_off);
            LSTC_LAUNCHED("k_orient_offsets(in)");
        }
    }
    if (d_sizes) *d_sizes = L.scal;
    if (d_bad) *d_bad = L.bad;
    return LSTC_OK;
}

extern "C" lstc_status lstc_clean_orient(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                                         uint32_t n, int compact, uint64_t *d_sym_off,
                                         uint32_t *d_sym_col, uint64_t *d_off, uint32_t *d_col,
                                         uint64_t *d_in_off, uint32_t *d_in_col, uint32_t *n_out,
                                         uint64_t *m_out, void *d_ws, size_t ws_bytes, lstc_stream s) {
    hipStream_t st = (hipStream_t)s;
    if (!n_out || !m_out) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: NULL size outputs");
    const uint64_t *d_sizes = nullptr;
    const uint32_t *d_bad = nullptr;
    if ((d_in_off == nullptr) != (d_in_col == nullptr))
        return fail(LSTC_ERR_INVALID, "lstc_clean_orient: give both in-list outputs or neither");
    LSTC_TRY(clean_orient_launch(d_src, d_dst, m_raw, n, compact, d_sym_off, d_sym_col, d_off, d_col, d_in_off,
                                 d_in_col, d_ws, ws_bytes, st, &d_sizes, &d_bad));
    uint64_t h[4];
    uint32_t bad = 0;
    LSTC_HIP(hipMemcpyAsync(h, d_sizes, sizeof(h), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    if (bad) return fail(LSTC_ERR_INVALID, "lstc_clean_orient: vertex id >= n in input");
    *n_out = (uint32_t)h[1];
    *m_out = h[2];
    return LSTC_OK;
}
