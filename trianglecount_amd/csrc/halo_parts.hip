// halo_parts.hip -- (b) the halo subgraphs of ALL k parts of a vertex partition at once.
//
// Same output per part as lstc_halo_extract (halo.hip), which follows
// create_subgraph_with_halo of preprocess_GPU/Halo_cleanned_preprocess/
// round_clean_subgraph.cu:357-902: P = {v : part[v] == p} ascending, H = N(P)\P
// (mode 0) or N+(P)\P (mode 1) ascending, local ids P block then H block, arcs
// by the global (degree, id) orientation, rows ascending.  The reference builds
// the parts one after another, each a full pass over the graph (the caller's
// loop over p, round_clean_subgraph.cu:904-976 / Halo_creation_preprocess.py);
// here every pass over the symmetric CSR serves all k parts:
//
//   membership : per vertex a k-bit mask of the parts whose local set holds it
//                (its own part, and every part of an in-side neighbour), pulled
//                from its neighbours: no atomics, one gather per arc;
//   local ids  : per part, tile counts + one scan over tiles + in-tile ranks;
//   rows       : per vertex, one walk over its arcs; lane p of the wave keeps
//                part p's counters (lengths) or write positions (fill), so an
//                arc kept by the orientation is tested against every part with
//                one ballot per part and the neighbour's 16-byte record
//                {mask, degree, part} is gathered once for all parts.
// A vertex v with part[v] >= k is owned by no part; it can still be in the halo
// of the parts of its neighbours, as lstc_halo_extract would put it there.
#include <vector>

#include "lstc_internal.h"
#include "wave.h"

namespace lstc {

static constexpr uint32_t kNoLoc = 0xFFFFFFFFu;
static constexpr uint32_t kMaxParts = 64;
static constexpr uint32_t kTileV = 1024;  // vertices per tile of the local-id scan (256 threads x 4)

// Per-vertex record, gathered once per arc.  The vertex's local ids in the parts
// of its mask sit together in lids[lbase ...] (ascending part order), so the ids
// of one neighbour in all parts come from one cache line.
struct __align__(16) VInfo {
    unsigned long long mem;    // bit p: the vertex lies in part p's local set (P_p or H_p)
    unsigned long long lbase;  // first of its popcount(mem) local ids in lids
    uint32_t deg;              // symmetric degree
    uint32_t part;             // part[v] (kNoLoc when >= k)
    uint32_t pad[2];
};

// Position of part p's local id among a vertex's ids (parts below p in its mask).
__device__ __forceinline__ uint32_t lid_rank(unsigned long long mem, uint32_t p) {
    return (uint32_t)__popcll(mem & ((1ull << p) - 1ull));
}

struct PartOut {             // device copy of lstc_halo_out
    uint64_t *off;
    uint32_t *col;
};

__device__ __forceinline__ bool keep_arc(uint32_t ds, uint32_t dd, uint32_t s, uint32_t d) {
    return ds < dd || (ds == dd && d < s);
}

__device__ __forceinline__ unsigned long long wave_or64(unsigned long long v) {
#pragma unroll
    for (uint32_t d = 32; d >= 1; d >>= 1) v |= gather64(v, lane_id() ^ d);
    return v;
}

__global__ void k_hp_init(const uint64_t *__restrict__ off, const uint32_t *__restrict__ part, uint32_t n,
                          uint32_t k, VInfo *__restrict__ info) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = part[v];
        VInfo x;
        x.mem = 0;
        x.lbase = 0;
        x.deg = (uint32_t)(off[v + 1] - off[v]);
        x.part = p < k ? p : kNoLoc;
        x.pad[0] = x.pad[1] = 0;
        info[v] = x;
    }
}

// mem[v] = own part | parts of the neighbours d that put v into their halo:
// every neighbour (mode 0) or the neighbours with an arc d -> v (mode 1).
__global__ __launch_bounds__(256) void k_hp_mask(const uint64_t *__restrict__ off, const uint32_t *__restrict__ col,
                                                 const uint32_t *__restrict__ part, uint32_t n, uint32_t k, int mode,
                                                 VInfo *__restrict__ info) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw) {
        const uint64_t b = off[v], e = off[v + 1];
        const uint32_t dv = (uint32_t)(e - b);
        unsigned long long m = 0;
        for (uint64_t j = b + lane; j < e; j += 64) {
            const uint32_t d = col[j];
            if (mode == 0) {  // the 4-byte part id is all a neighbour needs
                const uint32_t pd = part[d];
                if (pd < k) m |= 1ull << pd;
            } else {
                const VInfo x = info[d];
                if (x.part != kNoLoc && keep_arc(x.deg, dv, d, (uint32_t)v)) m |= 1ull << x.part;
            }
        }
        m = wave_or64(m);
        if (lane == 0) {
            const uint32_t p = info[v].part;
            info[v].mem = p == kNoLoc ? m : m | (1ull << p);
        }
    }
}

__global__ void k_hp_popc(const VInfo *__restrict__ info, uint32_t n, uint32_t *__restrict__ pc) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        pc[v] = (uint32_t)__popcll(info[v].mem);
}

__global__ void k_hp_lbase(const uint64_t *__restrict__ lb, uint32_t n, VInfo *__restrict__ info) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        info[v].lbase = lb[v];
}

// P and H counts of every part in each tile of kTileV vertices: cnt[tile][2p] = |P_p n tile|,
// cnt[tile][2p+1] = |H_p n tile|.
__global__ __launch_bounds__(256) void k_hp_tilecount(const VInfo *__restrict__ info, uint32_t n, uint32_t k,
                                                      uint32_t *__restrict__ cnt) {
    __shared__ uint32_t s_w[4][2 * kMaxParts];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t v0 = (uint64_t)blockIdx.x * kTileV + threadIdx.x * 4;
    unsigned long long mem[4];
    uint32_t pp[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool in = v0 + i < n;
        mem[i] = in ? info[v0 + i].mem : 0ull;
        pp[i] = in ? info[v0 + i].part : kNoLoc;
    }
    for (uint32_t p = 0; p < k; ++p) {
        uint32_t cp = 0, ch = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            cp += pp[i] == p ? 1u : 0u;
            ch += ((mem[i] >> p) & 1ull) && pp[i] != p ? 1u : 0u;
        }
        cp = wave_sum32(cp);
        ch = wave_sum32(ch);
        if (lane == 0) {
            s_w[wv][2 * p] = cp;
            s_w[wv][2 * p + 1] = ch;
        }
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < 2 * k; c += blockDim.x)
        cnt[(uint64_t)blockIdx.x * 2 * k + c] = s_w[0][c] + s_w[1][c] + s_w[2][c] + s_w[3][c];
}

// Exclusive scan down every column of cnt[tiles][2k] (one wave per column);
// tot[c] receives the column total (tot[2p] = |P_p|, tot[2p+1] = |H_p|).
__global__ __launch_bounds__(1024) void k_hp_tilescan(uint32_t *__restrict__ cnt, uint32_t tiles, uint32_t k,
                                                      uint64_t *__restrict__ tot) {
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t c = wv; c < 2 * k; c += blockDim.x >> 6) {
        uint32_t run = 0;
        for (uint32_t t0 = 0; t0 < tiles; t0 += 64) {
            const uint32_t t = t0 + lane;
            const uint32_t x = t < tiles ? cnt[(uint64_t)t * 2 * k + c] : 0u;
            const uint32_t incl = wave_incl_scan(x, lane);
            if (t < tiles) cnt[(uint64_t)t * 2 * k + c] = run + incl - x;
            run += read_lane(incl, 63);
        }
        if (lane == 0) tot[c] = run;
    }
}

// The local id of v in every part p of its mask (P block, then H block), at
// lids[lbase(v) + lid_rank(mem(v), p)].
__global__ __launch_bounds__(256) void k_hp_local(const VInfo *__restrict__ info, uint32_t n, uint32_t k,
                                                  const uint32_t *__restrict__ cnt, const uint64_t *__restrict__ tot,
                                                  uint32_t *__restrict__ lids) {
    __shared__ uint32_t s_w[4][2];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t v0 = (uint64_t)blockIdx.x * kTileV + threadIdx.x * 4;
    unsigned long long mem[4], lb[4];
    uint32_t pp[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool in = v0 + i < n;
        mem[i] = in ? info[v0 + i].mem : 0ull;
        lb[i] = in ? info[v0 + i].lbase : 0ull;
        pp[i] = in ? info[v0 + i].part : kNoLoc;
    }
    for (uint32_t p = 0; p < k; ++p) {
        uint32_t fp[4], fh[4], cp = 0, ch = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            fp[i] = pp[i] == p ? 1u : 0u;
            fh[i] = ((mem[i] >> p) & 1ull) && pp[i] != p ? 1u : 0u;
            cp += fp[i];
            ch += fh[i];
        }
        const uint32_t ip = wave_incl_scan(cp, lane), ih = wave_incl_scan(ch, lane);
        __syncthreads();  // s_w of the previous part has been read
        if (lane == 63) {
            s_w[wv][0] = ip;
            s_w[wv][1] = ih;
        }
        __syncthreads();
        uint32_t bp = 0, bh = 0;
        for (uint32_t w = 0; w < wv; ++w) {
            bp += s_w[w][0];
            bh += s_w[w][1];
        }
        uint32_t rp = (uint32_t)cnt[(uint64_t)blockIdx.x * 2 * k + 2 * p] + bp + ip - cp;
        uint32_t rh = (uint32_t)(tot[2 * p] + cnt[(uint64_t)blockIdx.x * 2 * k + 2 * p + 1]) + bh + ih - ch;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (v0 + i >= n) break;
            if (fp[i]) lids[lb[i] + lid_rank(mem[i], p)] = rp++;
            else if (fh[i]) lids[lb[i] + lid_rank(mem[i], p)] = rh++;
        }
    }
}

// l2g[p][local id of v in p] = v (launched once the sizes are checked against the caller's capacities).
__global__ void k_hp_l2g(const VInfo *__restrict__ info, const uint32_t *__restrict__ lids, uint32_t n,
                         uint32_t *const *__restrict__ l2g) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const VInfo x = info[v];
        uint32_t r = 0;
        for (unsigned long long rest = x.mem; rest; rest &= rest - 1, ++r)
            l2g[__builtin_ctzll(rest)][lids[x.lbase + r]] = (uint32_t)v;
    }
}

// One wave per vertex s; FILL = 0 counts each part's kept arcs of row s
// (rowlen[p][local row], and those into P_p: rowp[p][local row]); FILL = 1
// writes them.
// Lane p holds part p's state: counters, or the two write positions.
template <int FILL>
__global__ __launch_bounds__(256) void k_hp_rows(const uint64_t *__restrict__ off, const uint32_t *__restrict__ col,
                                                 uint32_t n, uint32_t k, const VInfo *__restrict__ info,
                                                 const uint32_t *__restrict__ lids, uint32_t *__restrict__ rowlen,
                                                 uint32_t *__restrict__ rowp, unsigned long long *__restrict__ m_tot,
                                                 const PartOut *__restrict__ outs) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    unsigned long long arcs = 0;  // lane p: arcs of part p over this wave's rows (FILL = 0)
    for (uint64_t s = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += nw) {
        const unsigned long long ms = read_lane64(info[s].mem, 0);  // wave-uniform
        if (ms == 0) continue;
        const uint64_t sbase = read_lane64(info[s].lbase, 0);
        const uint64_t b = off[s], e = off[s + 1];
        const uint32_t ds = (uint32_t)(e - b);
        const bool mine = lane < k && ((ms >> lane) & 1ull);
        const uint32_t ls = mine ? lids[sbase + lid_rank(ms, lane)] : 0u;  // s's local row in part `lane`
        uint32_t cnt = 0, cntp = 0;
        uint64_t posp = 0, posh = 0;
        uint32_t *ocol = nullptr;
        if (FILL && mine) {
            posp = outs[lane].off[ls];
            posh = posp + rowp[(uint64_t)lane * n + ls];
            ocol = outs[lane].col;
        }
        for (uint64_t j0 = b; j0 < e; j0 += 64) {
            const uint64_t j = j0 + lane;
            uint32_t d = 0;
            unsigned long long md = 0, dmem = 0, dbase = 0;
            uint32_t pd = kNoLoc;
            if (j < e) {
                d = col[j];
                const VInfo x = info[d];
                if (keep_arc(ds, x.deg, (uint32_t)s, d)) {
                    md = x.mem & ms;
                    dmem = x.mem;
                    dbase = x.lbase;
                    pd = x.part;
                }
            }
            // with at most 8 parts, the target's local ids in all its parts (side by side in
            // lids) are loaded together here rather than one after another in the part loop
            uint32_t l8[8];
            if (FILL && k <= 8) {
                const uint32_t np = (uint32_t)__popcll(md ? dmem : 0ull);
#pragma unroll
                for (uint32_t r = 0; r < 8; ++r) l8[r] = r < np ? lids[dbase + r] : 0u;
            }
            for (unsigned long long rest = ms; rest;) {
                const uint32_t p = (uint32_t)__builtin_ctzll(rest);
                rest &= rest - 1;
                const bool in = (md >> p) & 1ull, inp = in && pd == p;
                const uint64_t mP = ballot(inp), mA = ballot(in);
                if (!FILL) {
                    if (lane == p) {
                        cnt += (uint32_t)__popcll(mA);
                        cntp += (uint32_t)__popcll(mP);
                    }
                } else {
                    const uint64_t mH = mA & ~mP;
                    if (mA) {
                        const uint64_t bp = read_lane64(posp, p), bh = read_lane64(posh, p);
                        uint32_t *oc = (uint32_t *)read_lane64((uint64_t)ocol, p);
                        if (in) {
                            const uint32_t rk = lid_rank(dmem, p);
                            uint32_t ld;
                            if (k <= 8) {
                                ld = l8[0];
#pragma unroll
                                for (uint32_t r = 1; r < 8; ++r) ld = rk == r ? l8[r] : ld;
                            } else {
                                ld = lids[dbase + rk];
                            }
                            if (inp) oc[bp + mask_rank(mP)] = ld;
                            else oc[bh + mask_rank(mH)] = ld;
                        }
                        if (lane == p) {
                            posp += (uint64_t)__popcll(mP);
                            posh += (uint64_t)__popcll(mH);
                        }
                    }
                }
            }
        }
        if (!FILL && mine) {
            rowlen[(uint64_t)lane * n + ls] = cnt;
            rowp[(uint64_t)lane * n + ls] = cntp;
            arcs += cnt;
        }
    }
    if (!FILL && lane < k && arcs) atomicAdd(m_tot + lane, arcs);
}

__global__ void k_hp_last(const PartOut *__restrict__ outs, const uint64_t *__restrict__ tot, uint32_t k,
                          const unsigned long long *__restrict__ m_tot) {
    for (uint32_t p = threadIdx.x; p < k; p += blockDim.x) outs[p].off[tot[2 * p] + tot[2 * p + 1]] = m_tot[p];
}

struct HPLayout {
    VInfo *info;
    uint32_t *lids, *pc, *rowlen, *rowp, *cnt;
    uint64_t *lbase;
    uint64_t *tot;               // [2k]
    unsigned long long *m_tot;   // [k]
    PartOut *outs;               // [k]
    uint32_t **l2g;              // [k]
    void *scan_ws;
    size_t scan_bytes, bytes;
    uint32_t tiles;
};

static HPLayout hp_layout(void *ws, size_t ws_bytes, uint32_t n, uint32_t k) {
    Carver c(ws, ws_bytes);
    HPLayout L;
    const size_t nn = (size_t)n + 1, kk = k ? k : 1;
    L.tiles = (uint32_t)(((uint64_t)n + kTileV - 1) / kTileV);
    L.info = c.take<VInfo>(nn);
    L.lids = c.take<uint32_t>(kk * nn);  // at most k local ids per vertex
    L.pc = c.take<uint32_t>(nn);
    L.lbase = c.take<uint64_t>(nn);
    L.rowlen = c.take<uint32_t>(kk * nn);
    L.rowp = c.take<uint32_t>(kk * nn);
    L.cnt = c.take<uint32_t>((size_t)(L.tiles ? L.tiles : 1) * 2 * kk);
    L.tot = c.take<uint64_t>(2 * kk);
    L.m_tot = c.take<unsigned long long>(kk);
    L.outs = c.take<PartOut>(kk);
    L.l2g = c.take<uint32_t *>(kk);
    L.scan_bytes = scan_ws_bytes(nn);
    L.scan_ws = c.take<char>(L.scan_bytes);
    L.bytes = c.need();
    return L;
}

static lstc_status hp_check(const uint64_t *off, const uint32_t *col, uint32_t n, const uint32_t *part, uint32_t k,
                            int mode, const char *who) {
    if (n == 0xFFFFFFFFu) return fail(LSTC_ERR_INVALID, "%s: n must be < 2^32-1", who);
    if (k == 0 || k > kMaxParts) return fail(LSTC_ERR_INVALID, "%s: k = %u parts, 1..%u supported", who, k, kMaxParts);
    if (mode != 0 && mode != 1) return fail(LSTC_ERR_INVALID, "%s: mode must be 0 or 1", who);
    if (n && (!off || !part)) return fail(LSTC_ERR_INVALID, "%s: NULL input", who);
    (void)col;
    return LSTC_OK;
}

// membership, local ids and row lengths of every part; sizes stay on the device
// (tot: |P_p|, |H_p|; m_tot: arcs of part p).
static lstc_status hp_front(const uint64_t *off, const uint32_t *col, uint32_t n, const uint32_t *part, uint32_t k,
                            int mode, const HPLayout &L, hipStream_t st) {
    const int cus = device_cus();
    const uint32_t grid_n = (uint32_t)hmin<uint64_t>(((uint64_t)n + 255) / 256, (uint64_t)cus * 32);
    const uint32_t grid_w = (uint32_t)hmin<uint64_t>(((uint64_t)n + 3) / 4, (uint64_t)cus * 32);
    hipLaunchKernelGGL(k_hp_init, dim3(grid_n), dim3(256), 0, st, off, part, n, k, L.info);
    LSTC_LAUNCHED("k_hp_init");
    hipLaunchKernelGGL(k_hp_mask, dim3(grid_w), dim3(256), 0, st, off, col, part, n, k, mode, L.info);
    LSTC_LAUNCHED("k_hp_mask");
    hipLaunchKernelGGL(k_hp_popc, dim3(grid_n), dim3(256), 0, st, (const VInfo *)L.info, n, L.pc);
    LSTC_LAUNCHED("k_hp_popc");
    LSTC_TRY(scan_exclusive_u32_to_u64(L.pc, L.lbase, n, nullptr, L.scan_ws, L.scan_bytes, st));
    hipLaunchKernelGGL(k_hp_lbase, dim3(grid_n), dim3(256), 0, st, (const uint64_t *)L.lbase, n, L.info);
    LSTC_LAUNCHED("k_hp_lbase");
    hipLaunchKernelGGL(k_hp_tilecount, dim3(L.tiles), dim3(256), 0, st, L.info, n, k, L.cnt);
    LSTC_LAUNCHED("k_hp_tilecount");
    hipLaunchKernelGGL(k_hp_tilescan, dim3(1), dim3(1024), 0, st, L.cnt, L.tiles, k, L.tot);
    LSTC_LAUNCHED("k_hp_tilescan");
    hipLaunchKernelGGL(k_hp_local, dim3(L.tiles), dim3(256), 0, st, L.info, n, k, L.cnt, L.tot, L.lids);
    LSTC_LAUNCHED("k_hp_local");
    LSTC_TRY(fill_u64((uint64_t *)L.m_tot, 0, k, st));
    hipLaunchKernelGGL(k_hp_rows<0>, dim3(grid_w), dim3(256), 0, st, off, col, n, k, L.info, L.lids, L.rowlen, L.rowp,
                       L.m_tot, (const PartOut *)nullptr);
    LSTC_LAUNCHED("k_hp_rows<0>");
    return LSTC_OK;
}

static lstc_status hp_sizes_to_host(const HPLayout &L, uint32_t k, std::vector<uint64_t> &h, hipStream_t st) {
    h.assign(3 * (size_t)k, 0);
    LSTC_HIP(hipMemcpyAsync(h.data(), L.tot, 2 * k * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipMemcpyAsync(h.data() + 2 * k, L.m_tot, k * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipStreamSynchronize(st));
    return LSTC_OK;
}

}  // namespace lstc

using namespace lstc;

extern "C" lstc_status lstc_halo_parts_workspace(uint32_t n, uint32_t k, size_t *bytes) {
    if (!bytes) return fail(LSTC_ERR_INVALID, "lstc_halo_parts_workspace: NULL bytes");
    if (k == 0 || k > kMaxParts) return fail(LSTC_ERR_INVALID, "lstc_halo_parts_workspace: k = %u parts, 1..%u", k, kMaxParts);
    *bytes = hp_layout(nullptr, 0, n, k).bytes;
    return LSTC_OK;
}

extern "C" lstc_status lstc_halo_parts_sizes(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                             const uint32_t *d_part, uint32_t k, int mode, uint32_t *n_local,
                                             uint32_t *n_owned, uint64_t *m_local, void *d_ws, size_t ws_bytes,
                                             lstc_stream s) {
    hipStream_t st = (hipStream_t)s;
    LSTC_TRY(hp_check(d_sym_off, d_sym_col, n, d_part, k, mode, "lstc_halo_parts_sizes"));
    if (!n_local || !n_owned || !m_local) return fail(LSTC_ERR_INVALID, "lstc_halo_parts_sizes: NULL output");
    const HPLayout L = hp_layout(d_ws, ws_bytes, n, k);
    if (!d_ws || ws_bytes < L.bytes) return fail(LSTC_ERR_WORKSPACE, "lstc_halo_parts_sizes: workspace too small");
    if (n == 0) {
        for (uint32_t p = 0; p < k; ++p) { n_local[p] = n_owned[p] = 0; m_local[p] = 0; }
        return LSTC_OK;
    }
    LSTC_TRY(hp_front(d_sym_off, d_sym_col, n, d_part, k, mode, L, st));
    std::vector<uint64_t> h;
    LSTC_TRY(hp_sizes_to_host(L, k, h, st));
    for (uint32_t p = 0; p < k; ++p) {
        n_owned[p] = (uint32_t)h[2 * p];
        n_local[p] = (uint32_t)(h[2 * p] + h[2 * p + 1]);
        m_local[p] = h[2 * k + p];
    }
    return LSTC_OK;
}

// Front once, sizes to the host (n_local, n_owned, m_local), capacity checks,
// then offsets and columns of every part.  A part whose output is too small
// leaves every output untouched and fails with the sizes filled in.
static lstc_status hp_extract(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n, const uint32_t *d_part,
                              uint32_t k, int mode, const lstc_halo_out *outs, uint32_t *n_local, uint32_t *n_owned,
                              uint64_t *m_local, void *d_ws, size_t ws_bytes, hipStream_t st, const char *who) {
    LSTC_TRY(hp_check(d_sym_off, d_sym_col, n, d_part, k, mode, who));
    if (!outs) return fail(LSTC_ERR_INVALID, "%s: NULL outputs", who);
    const HPLayout L = hp_layout(d_ws, ws_bytes, n, k);
    if (!d_ws || ws_bytes < L.bytes) return fail(LSTC_ERR_WORKSPACE, "%s: workspace too small", who);
    bool with_l2g = false;
    for (uint32_t p = 0; p < k; ++p) {
        if (!outs[p].off) return fail(LSTC_ERR_INVALID, "%s: part %u has no offsets output", who, p);
        with_l2g = with_l2g || outs[p].l2g;
    }
    if (n == 0) {
        for (uint32_t p = 0; p < k; ++p) {
            n_local[p] = n_owned[p] = 0;
            m_local[p] = 0;
            LSTC_TRY(fill_u64(outs[p].off, 0, 1, st));
        }
        return LSTC_OK;
    }
    std::vector<PartOut> po(k);
    std::vector<uint32_t *> l2g(k);
    for (uint32_t p = 0; p < k; ++p) {
        po[p] = PartOut{outs[p].off, outs[p].col};
        l2g[p] = outs[p].l2g;
    }
    LSTC_HIP(hipMemcpyAsync(L.outs, po.data(), k * sizeof(PartOut), hipMemcpyHostToDevice, st));
    LSTC_HIP(hipMemcpyAsync(L.l2g, l2g.data(), k * sizeof(uint32_t *), hipMemcpyHostToDevice, st));
    LSTC_TRY(hp_front(d_sym_off, d_sym_col, n, d_part, k, mode, L, st));
    // sizes bound the writes into the caller's buffers
    std::vector<uint64_t> h;
    LSTC_TRY(hp_sizes_to_host(L, k, h, st));
    bool fits = true;
    for (uint32_t p = 0; p < k; ++p) {
        n_owned[p] = (uint32_t)h[2 * p];
        n_local[p] = (uint32_t)(h[2 * p] + h[2 * p + 1]);
        m_local[p] = h[2 * k + p];
        fits = fits && n_local[p] <= outs[p].n_local_cap && m_local[p] <= outs[p].m_local_cap;
    }
    for (uint32_t p = 0; p < k; ++p) {
        if (n_local[p] > outs[p].n_local_cap || m_local[p] > outs[p].m_local_cap)
            return fail(LSTC_ERR_INVALID, "%s: part %u output capacity too small (need %u rows, %llu arcs)", who, p,
                        n_local[p], (unsigned long long)m_local[p]);
        if (m_local[p] && !outs[p].col) return fail(LSTC_ERR_INVALID, "%s: part %u has no column output", who, p);
        if (outs[p].l2g == nullptr && with_l2g)
            return fail(LSTC_ERR_INVALID, "%s: give every part an l2g output or none", who);
    }
    (void)fits;
    for (uint32_t p = 0; p < k; ++p)
        if (n_local[p])
            LSTC_TRY(scan_exclusive_u32_to_u64(L.rowlen + (size_t)p * n, outs[p].off, n_local[p], nullptr, L.scan_ws,
                                               L.scan_bytes, st));
    hipLaunchKernelGGL(k_hp_last, dim3(1), dim3(64), 0, st, (const PartOut *)L.outs, L.tot, k,
                       (const unsigned long long *)L.m_tot);
    LSTC_LAUNCHED("k_hp_last");
    const int cus = device_cus();
    if (with_l2g) {
        const uint32_t grid = (uint32_t)hmin<uint64_t>(((uint64_t)n + 255) / 256, (uint64_t)cus * 32);
        hipLaunchKernelGGL(k_hp_l2g, dim3(grid), dim3(256), 0, st, (const VInfo *)L.info, (const uint32_t *)L.lids, n,
                           (uint32_t *const *)L.l2g);
        LSTC_LAUNCHED("k_hp_l2g");
    }
    const uint32_t grid_w = (uint32_t)hmin<uint64_t>(((uint64_t)n + 3) / 4, (uint64_t)cus * 32);
    hipLaunchKernelGGL(k_hp_rows<1>, dim3(grid_w), dim3(256), 0, st, d_sym_off, d_sym_col, n, k, L.info, L.lids,
                       L.rowlen, L.rowp, L.m_tot, (const PartOut *)L.outs);
    LSTC_LAUNCHED("k_hp_rows<1>");
    return LSTC_OK;
}

extern "C" lstc_status lstc_halo_parts_extract(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                               const uint32_t *d_part, uint32_t k, int mode, const lstc_halo_out *outs,
                                               void *d_ws, size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_halo_parts_extract");
    if (k == 0 || k > kMaxParts) return fail(LSTC_ERR_INVALID, "lstc_halo_parts_extract: k = %u parts, 1..%u", k, kMaxParts);
    std::vector<uint32_t> nl(k), no(k);
    std::vector<uint64_t> ml(k);
    return hp_extract(d_sym_off, d_sym_col, n, d_part, k, mode, outs, nl.data(), no.data(), ml.data(), d_ws, ws_bytes,
                      (hipStream_t)s, "lstc_halo_parts_extract");
}

extern "C" lstc_status lstc_halo_parts_build(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                             const uint32_t *d_part, uint32_t k, int mode, const lstc_halo_out *outs,
                                             uint32_t *n_local, uint32_t *n_owned, uint64_t *m_local, void *d_ws,
                                             size_t ws_bytes, lstc_stream s) {
    TraceRange range_("lstc_halo_parts_build");
    if (!n_local || !n_owned || !m_local) return fail(LSTC_ERR_INVALID, "lstc_halo_parts_build: NULL size outputs");
    return hp_extract(d_sym_off, d_sym_col, n, d_part, k, mode, outs, n_local, n_owned, m_local, d_ws, ws_bytes,
                      (hipStream_t)s, "lstc_halo_parts_build");
}
