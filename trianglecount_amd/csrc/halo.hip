// halo.hip -- (b) halo-subgraph extraction for one part of a k-way vertex partition.
//
// Replaces create_subgraph_with_halo of preprocess_GPU/Halo_cleanned_preprocess/
// round_clean_subgraph.cu:357-902 (find_halo_nodes1 + merge sort + unique,
// Find_size + InclusiveSum + Convert, sorting_col_idx):
//   P  = {v : part[v] == p} in ascending id                    (:393-397)
//   H  = N(P) \ P (mode 0) or N+(P) \ P (mode 1), ascending     (:424, :468-495)
//   local ids: P block then H block                             (:527-530, :600-602)
//   arcs: the global (degree, id) orientation restricted to P u H (:160-272)
//   rows ascending by local id                                  (:323-346, :656)
// Mode 1 keeps exactly the vertices an owned triangle can touch (the intent of
// the pruning loop :671-838, whose later rounds are defective in the reference).
//
// GPU plan: a membership byte map replaces the per-neighbour binary searches and
// the sort/unique of halo candidates; two prefix sums give monotone local ids;
// a wave per local row counts, and then writes, its kept arcs.  Because every P
// local id is below every H local id, a row in local order is its P-targets in
// global order followed by its H-targets in global order: the wave writes the
// two runs at their final positions with ballot ranks, no per-row sort.
#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ bool keep_arc(uint64_t ds, uint64_t dd, uint32_t s, uint32_t d) {
    return ds < dd || (ds == dd && d < s);
}

// mark[d] = 1 for every d in N(P)\P (mode 0) or N+(P)\P (mode 1).  One wave per row.
__global__ __launch_bounds__(256) void k_halo_mark(const uint64_t *__restrict__ off,
                                                   const uint32_t *__restrict__ col, uint32_t n,
                                                   const uint32_t *__restrict__ part, uint32_t p,
                                                   int mode, uint8_t *__restrict__ mark) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw) {
        if (part[v] != p) continue;
        const uint64_t b = off[v], e = off[v + 1];
        const uint64_t dv = e - b;
        for (uint64_t j = b + lane; j < e; j += 64) {
            const uint32_t d = col[j];
            if (part[d] == p) continue;
            if (mode == 1 && !keep_arc(dv, off[d + 1] - off[d], (uint32_t)v, d)) continue;
            mark[d] = 1;
        }
    }
}

__global__ void k_halo_flags(const uint32_t *__restrict__ part, uint32_t p, uint32_t n,
                             const uint8_t *__restrict__ mark, uint32_t *__restrict__ fP,
                             uint32_t *__restrict__ fH) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const bool inP = part[v] == p;
        fP[v] = inP ? 1u : 0u;
        fH[v] = (!inP && mark[v]) ? 1u : 0u;
    }
}

// local[v] and l2g[local] from the two exclusive scans; totals in tot[0] (|P|), tot[1] (|H|).
__global__ void k_halo_local(const uint32_t *__restrict__ fP, const uint32_t *__restrict__ fH,
                             const uint64_t *__restrict__ sP, const uint64_t *__restrict__ sH,
                             const uint64_t *__restrict__ tot, uint32_t n, uint32_t *__restrict__ local,
                             uint32_t *__restrict__ l2g) {
    const uint64_t np = tot[0];
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t id = kNone;
        if (fP[v]) id = (uint32_t)sP[v];
        else if (fH[v]) id = (uint32_t)(np + sH[v]);
        local[v] = id;
        if (id != kNone) l2g[id] = (uint32_t)v;
    }
}

// Per local row: kept arc count (and how many point into P).  One wave per row.
__global__ __launch_bounds__(256) void k_halo_rowlen(const uint64_t *__restrict__ off,
                                                     const uint32_t *__restrict__ col,
                                                     const uint32_t *__restrict__ local,
                                                     const uint32_t *__restrict__ l2g,
                                                     const uint64_t *__restrict__ tot,
                                                     uint32_t *__restrict__ rowlen,
                                                     uint32_t *__restrict__ rowP) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t np = tot[0], nl = tot[0] + tot[1];
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t i = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < nl; i += nw) {
        const uint32_t s = l2g[i];
        const uint64_t b = off[s], e = off[s + 1], ds = e - b;
        uint32_t cnt = 0, cntP = 0;
        for (uint64_t j0 = b; j0 < e; j0 += 64) {
            const uint64_t j = j0 + lane;
            bool keep = false, isP = false;
            if (j < e) {
                const uint32_t d = col[j];
                const uint32_t ld = local[d];
                if (ld != kNone && keep_arc(ds, off[d + 1] - off[d], s, d)) {
                    keep = true;
                    isP = ld < np;
                }
            }
            cnt += (uint32_t)__popcll(ballot(keep));
            cntP += (uint32_t)__popcll(ballot(isP));
        }
        if (lane == 0) {
            rowlen[i] = cnt;
            rowP[i] = cntP;
        }
    }
}

__global__ __launch_bounds__(256) void k_halo_fill(const uint64_t *__restrict__ off,
                                                   const uint32_t *__restrict__ col,
                                                   const uint32_t *__restrict__ local,
                                                   const uint32_t *__restrict__ l2g,
                                                   const uint64_t *__restrict__ tot,
                                                   const uint64_t *__restrict__ out_off,
                                                   const uint32_t *__restrict__ rowP,
                                                   uint32_t *__restrict__ out_col) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t np = tot[0], nl = tot[0] + tot[1];
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t i = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < nl; i += nw) {
        const uint32_t s = l2g[i];
        const uint64_t b = off[s], e = off[s + 1], ds = e - b;
        uint64_t posP = out_off[i], posH = out_off[i] + rowP[i];
        for (uint64_t j0 = b; j0 < e; j0 += 64) {
            const uint64_t j = j0 + lane;
            bool keep = false, isP = false;
            uint32_t ld = kNone;
            if (j < e) {
                const uint32_t d = col[j];
                ld = local[d];
                if (ld != kNone && keep_arc(ds, off[d + 1] - off[d], s, d)) {
                    keep = true;
                    isP = ld < np;
                }
            }
            const uint64_t mP = ballot(isP), mH = ballot(keep && !isP);
            if (isP) out_col[posP + mask_rank(mP)] = ld;
            else if (keep) out_col[posH + mask_rank(mH)] = ld;
            posP += (uint64_t)__popcll(mP);
            posH += (uint64_t)__popcll(mH);
        }
    }
}

__global__ void k_halo_final(const uint64_t *__restrict__ tot, uint64_t *__restrict__ out_off,
                             const uint64_t *__restrict__ d_m) {
    if (threadIdx.x == 0 && blockIdx.x == 0) out_off[tot[0] + tot[1]] = *d_m;
}

struct HaloLayout {
    uint8_t *mark;
    uint32_t *fP, *fH, *local, *l2g, *rowlen, *rowP;
    uint64_t *sP, *sH, *tot;  // tot[0]=|P| tot[1]=|H| tot[2]=m_local
    uint64_t *off_tmp;
    void *scan_ws;
    size_t scan_bytes, bytes;
};

static HaloLayout halo_layout(void *ws, size_t ws_bytes, uint32_t n) {
    Carver c(ws, ws_bytes);
    HaloLayout L;
    const size_t nn = (size_t)n + 1;
    L.mark = c.take<uint8_t>(nn);
    L.fP = c.take<uint32_t>(nn);
    L.fH = c.take<uint32_t>(nn);
    L.local = c.take<uint32_t>(nn);
    L.l2g = c.take<uint32_t>(nn);
    L.rowlen = c.take<uint32_t>(nn);
    L.rowP = c.take<uint32_t>(nn);
    L.sP = c.take<uint64_t>(nn);
    L.sH = c.take<uint64_t>(nn);
    L.off_tmp = c.take<uint64_t>(nn + 1);
    L.tot = c.take<uint64_t>(8);
    L.scan_bytes = scan_ws_bytes(nn);
    L.scan_ws = c.take<char>(L.scan_bytes);
    L.bytes = c.need();
    return L;
}

// Runs everything up to the row offsets; out_off receives n_local+1 offsets
// (into L.off_tmp when the caller only wants sizes).
static lstc_status halo_front(const uint64_t *off, const uint32_t *col, uint32_t n, const uint32_t *part,
                              uint32_t p, int mode, HaloLayout &L, uint64_t *out_off, hipStream_t st) {
    const int cus = device_cus();
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n + 255) / 256 + 1, (uint64_t)cus * 32);
    const uint32_t grid_w = (uint32_t)hmin<uint64_t>(((uint64_t)n + 3) / 4 + 1, (uint64_t)cus * 16);
    LSTC_TRY(fill_u8(L.mark, 0, (uint64_t)n + 1, st));
    LSTC_TRY(fill_u64(L.tot, 0, 8, st));
    hipLaunchKernelGGL(k_halo_mark, dim3(grid_w), dim3(256), 0, st, off, col, n, part, p, mode, L.mark);
    LSTC_LAUNCHED("k_halo_mark");
    hipLaunchKernelGGL(k_halo_flags, dim3(grid_n), dim3(256), 0, st, part, p, n, L.mark, L.fP, L.fH);
    LSTC_LAUNCHED("k_halo_flags");
    LSTC_TRY(scan_exclusive_u32_to_u64(L.fP, L.sP, n, L.tot + 0, L.scan_ws, L.scan_bytes, st));
    LSTC_TRY(scan_exclusive_u32_to_u64(L.fH, L.sH, n, L.tot + 1, L.scan_ws, L.scan_bytes, st));
    hipLaunchKernelGGL(k_halo_local, dim3(grid_n), dim3(256), 0, st, L.fP, L.fH, L.sP, L.sH, L.tot, n,
                       L.local, L.l2g);
    LSTC_LAUNCHED("k_halo_local");
    // rows beyond n_local keep rowlen 0 so the scan over n entries is exact
    LSTC_TRY(fill_u32(L.rowlen, 0, (uint64_t)n + 1, st));
    hipLaunchKernelGGL(k_halo_rowlen, dim3(grid_w), dim3(256), 0, st, off, col, L.local, L.l2g, L.tot,
                       L.rowlen, L.rowP);
    LSTC_LAUNCHED("k_halo_rowlen");
    LSTC_TRY(scan_exclusive_u32_to_u64(L.rowlen, out_off, n, L.tot + 2, L.scan_ws, L.scan_bytes, st));
    return LSTC_OK;
}

static lstc_status halo_check_args(const uint64_t *off, const uint32_t *col, uint32_t n,
                                   const uint32_t *part) {
    if (n == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "halo: n must be < 2^32-1");
    if (n && (!off || !part)) return fail(LSTC_ERR_INVALID, "halo: NULL input");
    (void)col;
    return LSTC_OK;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_halo_workspace(uint32_t n, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_halo_workspace: NULL bytes");
    *bytes = halo_layout(nullptr, 0, n).bytes;
    return LSTC_OK;
}

extern "C" lstc_status lstc_halo_sizes(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                       const uint32_t *d_part, uint32_t p, int mode, uint32_t *n_local,
                                       uint32_t *n_owned, uint64_t *m_local, void *d_ws, size_t ws_bytes,
                                       lstc_stream s) {
    TraceRange range_("lstc_halo_sizes");
    hipStream_t st = (hipStream_t)s;
    LSTC_TRY(halo_check_args(d_sym_off, d_sym_col, n, d_part));
    if (!n_local || !n_owned || !m_local) return fail(LSTC_ERR_INVALID, "lstc_halo_sizes: NULL output");
    if (mode != 0 && mode != 1) return fail(LSTC_ERR_INVALID, "lstc_halo_sizes: mode must be 0 or 1");
    HaloLayout L = halo_layout(d_ws, ws_bytes, n);
    if (!d_ws || ws_bytes < L.bytes) return fail(LSTC_ERR_WORKSPACE, "lstc_halo_sizes: workspace too small");
    if (n == 0) { *n_local = *n_owned = 0; *m_local = 0; return LSTC_OK; }
    LSTC_TRY(halo_front(d_sym_off, d_sym_col, n, d_part, p, mode, L, L.off_tmp, st));
    uint64_t h[3];
    LSTC_HIP(hipMemcpyAsync(h, L.tot, sizeof(h), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    *n_owned = (uint32_t)h[0];
    *n_local = (uint32_t)(h[0] + h[1]);
    *m_local = h[2];
    return LSTC_OK;
}

extern "C" lstc_status lstc_halo_extract(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                         const uint32_t *d_part, uint32_t p, int mode, uint64_t *d_off_out,
                                         uint32_t *d_col_out, uint32_t *d_l2g_out, uint32_t n_local_cap,
                                         uint64_t m_local_cap, void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_halo_extract");
    hipStream_t st = (hipStream_t)s;
    LSTC_TRY(halo_check_args(d_sym_off, d_sym_col, n, d_part));
    if (mode != 0 && mode != 1) return fail(LSTC_ERR_INVALID, "lstc_halo_extract: mode must be 0 or 1");
    if (!d_off_out) return fail(LSTC_ERR_INVALID, "lstc_halo_extract: NULL offsets output");
    HaloLayout L = halo_layout(d_ws, ws_bytes, n);
    if (!d_ws || ws_bytes < L.bytes) return fail(LSTC_ERR_WORKSPACE, "lstc_halo_extract: workspace too small");
    if (n == 0) return fill_u64(d_off_out, 0, 1, st);
    LSTC_TRY(halo_front(d_sym_off, d_sym_col, n, d_part, p, mode, L, L.off_tmp, st));
    // sizes are needed to bound the writes into caller buffers
    uint64_t h[3];
    LSTC_HIP(hipMemcpyAsync(h, L.tot, sizeof(h), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    const uint64_t nl = h[0] + h[1];
    if (nl > n_local_cap || h[2] > m_local_cap)
        return fail(LSTC_ERR_INVALID, "lstc_halo_extract: output capacity too small (need %llu rows, %llu arcs)",
                    (unsigned long long)nl, (unsigned long long)h[2]);
    if (h[2] && !d_col_out) return fail(LSTC_ERR_INVALID, "lstc_halo_extract: NULL column output");
    LSTC_HIP(hipMemcpyAsync(d_off_out, L.off_tmp, (nl + 1) * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_halo_final, dim3(1), dim3(64), 0, st, L.tot, d_off_out, L.tot + 2);
    LSTC_LAUNCHED("k_halo_final");
    const int cus = device_cus();
    const uint32_t grid_w = (uint32_t)hmin<uint64_t>((nl + 3) / 4 + 1, (uint64_t)cus * 16);
    if (h[2]) {
        hipLaunchKernelGGL(k_halo_fill, dim3(grid_w), dim3(256), 0, st, d_sym_off, d_sym_col, L.local, L.l2g,
                           L.tot, (const uint64_t *)d_off_out, L.rowP, d_col_out);
        LSTC_LAUNCHED("k_halo_fill");
    }
    if (d_l2g_out && nl)
        LSTC_HIP(hipMemcpyAsync(d_l2g_out, L.l2g, nl * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    return LSTC_OK;
}
