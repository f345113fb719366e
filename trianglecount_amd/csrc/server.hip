// server.hip -- continuously batched count serving with an HBM budget.
//
// The reference has no serving path: every graph is a separate process run
// (README.md:57-91).  Here a server owns one device slab of `budget` bytes and a
// worker thread that runs the loop
//   admit   : FIFO requests whose working set fits the free slab space get
//             memory and their preparation (seeded generation + clean/orient +
//             in-lists + segment words) enqueued on one of kPrepStreams prep
//             streams, round robin, so a large request does not hold up the
//             small ones behind it.  Generated requests waiting together are
//             prepared as one group (up to kGroupMax): each generates into its
//             own id range of one union edge list, the union goes through one
//             clean/orient pipeline (graphs never share an edge, so cleaning and
//             orientation act on each alone), and two kernels copy every
//             graph's rows back into its own CSR with local ids -- one launch
//             sequence for the group instead of ~45 launches per request, which
//             is what bounds a host thread under load;
//   collect : prepared groups (prep event done) release their transient
//             memory, have their out-of-range flag checked and their requests
//             wait for counting; each CSR stays resident;
//   launch  : while one of the kSlots batch slots is free, every prepared
//             request is counted by ONE persistent batched launch
//             (count_engine.hip) on the count stream, results copied to pinned
//             host memory behind an event;
//   finish  : when a slot's event completes, counts and latencies are
//             published, resident memory is freed and waiters are woken.
// Requests that arrive while batches are on the GPU are prepared concurrently
// and join the next launch, so the batch grows with the load (continuous
// batching).  With two slots the next batch is already queued behind the one
// that runs, so the device does not idle while the host notices a finished
// batch.  The loop is a host thread on purpose: nothing on the device waits on
// host flags, so every kernel runs to completion on its own and a lost or
// stalled host never leaves a spinning kernel on the GPU.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "count_engine.h"

namespace lstc {

using Clock = std::chrono::steady_clock;

// First-fit sub-allocator over one device slab (offsets aligned to 256 B).
class SlabPool {
   public:
    void init(char *base, size_t bytes) {
        base_ = base;
        free_.clear();
        free_[0] = bytes;
        total_ = bytes;
        used_ = 0;
    }
    char *alloc(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                sizes_[off] = bytes;
                return base_ + off;
            }
        }
        return nullptr;
    }
    void release(char *p) {
        if (!p) return;
        const size_t off = (size_t)(p - base_);
        auto s = sizes_.find(off);
        if (s == sizes_.end()) return;
        size_t bytes = s->second;
        sizes_.erase(s);
        used_ -= bytes;
        size_t start = off;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                bytes += prev->second;
                free_.erase(prev);
            }
        }
        next = free_.lower_bound(start + bytes);
        if (next != free_.end() && next->first == start + bytes) {
            bytes += next->second;
            free_.erase(next);
        }
        free_[start] = bytes;
    }
    size_t total() const { return total_; }
    size_t used() const { return used_.load(std::memory_order_relaxed); }  // read by other threads (stats)

   private:
    char *base_ = nullptr;
    std::map<size_t, size_t> free_;
    std::unordered_map<size_t, size_t> sizes_;
    size_t total_ = 0;
    std::atomic<size_t> used_{0};
};

struct Req {
    uint64_t id;
    lstc_request spec;
    Clock::time_point t_submit;
    // device memory
    char *resident = nullptr;
    uint64_t *off = nullptr;
    uint32_t *col = nullptr;
    uint64_t *in_off = nullptr;  // target-side in-lists (null for caller CSRs)
    uint32_t *in_col = nullptr;
    uint64_t *col_seg = nullptr;  // their segment words
    uint64_t *in_seg = nullptr;
    uint32_t n_rows = 0;
    uint64_t m_cap = 0;
};

// Generated requests prepared together (one or more).
struct PrepGroup {
    std::vector<Req *> reqs;
    char *transient = nullptr;  // raw edges, cleaning workspace (and for several graphs their union's CSR)
    hipEvent_t done = nullptr;
    const uint32_t *d_bad = nullptr;
    int flag_slot = -1;  // pinned host word that receives *d_bad at the end of the preparation
};

struct Result {
    lstc_status status;
    uint64_t count;
    uint64_t edges;
    double latency_s;
};

// edges[g] = off[n_rows] of every graph of a batch (undirected edges after cleaning)
__global__ void k_gather_edges(const GraphDev *__restrict__ desc, uint32_t G, unsigned long long *__restrict__ out) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
        out[g] = desc[g].off[desc[g].n_rows];
}

__host__ __device__ inline size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

// A request's resident block: oriented CSR (offsets at 0), its in-lists, then
// the segment words of both (byte offsets; arrays sized for m_raw arcs).
struct ResLayout {
    size_t col, in_off, in_col, col_seg, in_seg, bytes;
};
__host__ __device__ inline ResLayout res_layout(uint32_t n, uint64_t m_raw) {
    const size_t mr = m_raw ? m_raw : 1;
    const size_t offb = round256(8 * ((size_t)n + 1)), half = offb + round256(4 * mr);
    return ResLayout{offb, half, half + offb, 2 * half, 2 * half + round256(8 * mr), 2 * half + 2 * round256(8 * mr)};
}

static constexpr int kPrepStreams = 3;  // preparation streams (round robin over groups)
static constexpr uint32_t kGroupMax = 64;            // requests prepared together at most
static constexpr size_t kGroupsInFlight = 2 * kPrepStreams;  // groups in preparation at once: under load
                                                             // the requests behind them form larger groups
static constexpr uint64_t kGroupRawMax = 1ull << 25;  // their raw edges at most
static constexpr uint64_t kGroupRowsMax = 1ull << 28; // their vertices at most

// Where a group's graphs go: graph g holds union ids [base[g], base[g + 1])
// and its resident block starts at slab offset res[g] (sized for mcap[g] raw
// edges).  Passed by value as a kernel argument.
struct GroupTable {
    uint32_t G;
    uint32_t base[kGroupMax + 1];
    uint64_t res[kGroupMax];
    uint64_t mcap[kGroupMax];
};

// The graph g with base[g] <= x < base[g + 1] (x below base[G]) for the bounds
// b[0..G]; ascending.
__device__ __forceinline__ uint32_t group_find(const uint64_t *b, uint32_t G, uint64_t x) {
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (b[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// Rows of a group's union into each graph: offsets relative to the graph's
// first arc (and first in-list entry), the closing offset by the graph's last row.
__global__ void k_group_rows(GroupTable T, char *__restrict__ slab, const uint64_t *__restrict__ uoff,
                             const uint64_t *__restrict__ uin_off) {
    __shared__ uint64_t s_base[kGroupMax + 1];
    for (uint32_t g = threadIdx.x; g <= T.G; g += blockDim.x) s_base[g] = T.base[g];
    __syncthreads();
    const uint32_t N = T.base[T.G];
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const uint32_t g = group_find(s_base, T.G, v);
        const uint32_t b = T.base[g], n = T.base[g + 1] - b, i = v - b;
        const ResLayout L = res_layout(n, T.mcap[g]);
        char *blk = slab + T.res[g];
        uint64_t *off = (uint64_t *)blk, *in_off = (uint64_t *)(blk + L.in_off);
        const uint64_t a0 = uoff[b], i0 = uin_off[b];
        off[i] = uoff[v] - a0;
        in_off[i] = uin_off[v] - i0;
        if (i + 1 == n) {
            off[n] = uoff[v + 1] - a0;
            in_off[n] = uin_off[v + 1] - i0;
        }
    }
}

// Arcs and in-list entries of the union into each graph: ids minus the graph's
// first id, segment-word starts minus its first arc (lengths stay).
__global__ void k_group_arcs(GroupTable T, char *__restrict__ slab, const uint64_t *__restrict__ uoff,
                             const uint32_t *__restrict__ ucol, const uint64_t *__restrict__ useg,
                             const uint64_t *__restrict__ uin_off, const uint32_t *__restrict__ uin_col,
                             const uint64_t *__restrict__ uin_seg) {
    __shared__ uint64_t s_arc[kGroupMax + 1], s_in[kGroupMax + 1];
    for (uint32_t g = threadIdx.x; g <= T.G; g += blockDim.x) {
        s_arc[g] = uoff[T.base[g]];
        s_in[g] = uin_off[T.base[g]];
    }
    __syncthreads();
    const uint64_t M = s_arc[T.G], MI = s_in[T.G];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < M; p += stride) {
        const uint32_t g = group_find(s_arc, T.G, p);
        const ResLayout L = res_layout(T.base[g + 1] - T.base[g], T.mcap[g]);
        char *blk = slab + T.res[g];
        const uint64_t j = p - s_arc[g];
        ((uint32_t *)(blk + L.col))[j] = ucol[p] - T.base[g];
        ((uint64_t *)(blk + L.col_seg))[j] = useg[p] - s_arc[g];
    }
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < MI; q += stride) {
        const uint32_t g = group_find(s_in, T.G, q);
        const ResLayout L = res_layout(T.base[g + 1] - T.base[g], T.mcap[g]);
        char *blk = slab + T.res[g];
        const uint64_t j = q - s_in[g];
        ((uint32_t *)(blk + L.in_col))[j] = uin_col[q] - T.base[g];
        ((uint64_t *)(blk + L.in_seg))[j] = uin_seg[q] - s_arc[g];
    }
}
static constexpr int kSlots = 2;        // batches that may be queued on the count stream at once
static constexpr int kFlagSlots = 4096; // requests in preparation at once (pinned flag words)

// One batch in flight: its share of the reserve and its pinned staging.
struct BatchSlot {
    char *reserve = nullptr;
    GraphDev *h_desc = nullptr;
    uint64_t *h_base = nullptr;
    unsigned long long *h_counts = nullptr;
    unsigned long long *h_edges = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Req *> batch;  // empty = free
};

}  // namespace lstc

using namespace lstc;

struct lstc_server {
    int device = 0;
    size_t budget = 0;
    uint32_t max_batch = 0;
    lstc_tune tune{};
    bool has_tune = false;
    char *slab = nullptr;
    SlabPool pool;             // request working sets (generation, cleaning, resident CSRs)
    size_t reserve_bytes = 0;  // per batch slot: descriptors, counts and work queue (outside the pool)
    hipStream_t prep_stream[kPrepStreams] = {};
    hipStream_t count_stream = nullptr;
    uint32_t next_prep = 0;
    BatchSlot slot[kSlots];
    uint32_t next_slot = 0;    // the slot the next launch takes (slots complete in launch order)
    uint32_t *h_flags = nullptr;  // pinned out-of-range flags of groups in preparation
    std::vector<int> free_flags;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> submitted;          // guarded by mu
    std::unordered_set<uint64_t> queued, cancelled;  // guarded by mu: not yet admitted / asked to cancel
    std::unordered_map<uint64_t, Result> results;  // guarded by mu
    std::vector<double> latencies;        // guarded by mu
    uint64_t next_id = 1, n_submitted = 0, n_rejected = 0, n_completed = 0, n_batches = 0, edges_done = 0;
    uint64_t triangles_done = 0;
    uint64_t max_batch_seen = 0;
    uint64_t n_groups = 0, max_group_seen = 0;  // preparation groups
    bool stop = false;
    lstc_status fatal = LSTC_OK;
    std::thread worker;

    // worker-private state
    std::deque<Req *> pending, ready;
    std::vector<PrepGroup *> preparing;
};

namespace lstc {

static size_t batch_head_bytes(uint32_t G);

static size_t prep_bytes(uint32_t n, uint64_t m_raw, size_t *ws_out) {
    size_t ws = 0;
    lstc_clean_orient_workspace(n, m_raw, &ws);
    *ws_out = ws;
    return round256(8 * m_raw) + ws;  // raw src/dst + clean workspace
}
// resident per request: oriented CSR, its target-side in-lists (both at most
// m_raw arcs) and their segment words (split count)
static size_t resident_bytes(uint32_t n, uint64_t m_raw) { return res_layout(n, m_raw).bytes; }
// transient of a group of several graphs: raw union edges, the union's cleaning
// workspace and its CSR (with in-lists and segment words)
static size_t group_bytes(uint32_t N, uint64_t M, size_t *ws_out) {
    size_t ws = 0;
    lstc_clean_orient_workspace(N, M, &ws);
    *ws_out = ws;
    return round256(8 * M) + round256(ws) + res_layout(N, M).bytes;
}
static uint32_t request_rows(const lstc_request &r) {
    if (r.kind == LSTC_REQ_RMAT) return 1u << r.scale;
    if (r.kind == LSTC_REQ_ER) return r.n;
    return r.n_rows;
}

// Admit the head of `pending` while memory allows; returns false on a HIP error.
// A request asked to be cancelled is dropped here, before it takes memory.
static lstc_status admit(lstc_server *S) {
    while (!S->pending.empty()) {
        Req *r = S->pending.front();
        const bool csr = r->spec.kind == LSTC_REQ_CSR;
        {
            // a cancel is honoured exactly when it lands before the request leaves
            // `queued`, which happens under the same lock as this check
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->cancelled.erase(r->id)) {
                S->queued.erase(r->id);
                S->results[r->id] = Result{LSTC_ERR_CANCELLED, 0, 0,
                                           std::chrono::duration<double>(Clock::now() - r->t_submit).count()};
                S->n_completed++;
                S->pending.pop_front();
                delete r;
                S->cv_done.notify_all();
                continue;
            }
            if (csr) S->queued.erase(r->id);
        }
        if (csr) {  // caller-resident CSR: nothing to prepare
            r->off = (uint64_t *)r->spec.off;
            r->col = (uint32_t *)r->spec.col;
            r->n_rows = r->spec.n_rows;
            r->m_cap = r->spec.m;
            S->pending.pop_front();
            S->ready.push_back(r);
            continue;
        }
        if (S->free_flags.empty() || S->preparing.size() >= kGroupsInFlight) break;  // enough in preparation for now
        // the group: the head and the generated requests right behind it (up to
        // the first caller CSR or cancelled request), within the group limits
        uint32_t G = 1;
        uint64_t N = request_rows(r->spec), M = r->spec.m_raw;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            while (G < kGroupMax && G < S->pending.size()) {
                const Req *q = S->pending[G];
                if (q->spec.kind == LSTC_REQ_CSR || S->cancelled.count(q->id)) break;
                const uint64_t nq = request_rows(q->spec);
                if (N + nq > kGroupRowsMax || M + q->spec.m_raw > kGroupRawMax) break;
                N += nq;
                M += q->spec.m_raw;
                ++G;
            }
        }
        // memory for the group, halved until it fits (one request alone, or wait)
        char *t = nullptr;
        std::vector<char *> res;
        size_t ws = 0;
        for (;;) {
            t = S->pool.alloc(G == 1 ? prep_bytes((uint32_t)N, M, &ws) : group_bytes((uint32_t)N, M, &ws));
            bool ok = t != nullptr;
            for (uint32_t g = 0; ok && g < G; ++g) {
                const Req *q = S->pending[g];
                char *blk = S->pool.alloc(resident_bytes(request_rows(q->spec), q->spec.m_raw));
                if (blk) res.push_back(blk); else ok = false;
            }
            if (ok) break;
            for (char *blk : res) S->pool.release(blk);
            res.clear();
            S->pool.release(t);
            t = nullptr;
            if (G == 1) break;
            G /= 2;
            N = M = 0;
            for (uint32_t g = 0; g < G; ++g) {
                N += request_rows(S->pending[g]->spec);
                M += S->pending[g]->spec.m_raw;
            }
        }
        if (!t) break;
        // commit: the requests leave `queued` (a later cancel is refused) unless
        // a cancel arrived meanwhile, where the group stops before that request
        uint32_t keep = G;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (uint32_t g = 0; g < G; ++g)
                if (S->cancelled.count(S->pending[g]->id)) {
                    keep = g;
                    break;
                }
            for (uint32_t g = 0; g < keep; ++g) S->queued.erase(S->pending[g]->id);
            if (keep) {
                S->n_groups++;
                S->max_group_seen = std::max<uint64_t>(S->max_group_seen, keep);
            }
        }
        if (keep < G) {
            for (uint32_t g = keep; g < G; ++g) S->pool.release(res[g]);
            res.resize(keep);
            if (keep == 0) {  // the head was cancelled: the top of the loop drops it
                S->pool.release(t);
                continue;
            }
            // the transient block and workspace size stay those of the larger group (they cover the smaller one)
            G = keep;
            N = M = 0;
            for (uint32_t g = 0; g < G; ++g) {
                N += request_rows(S->pending[g]->spec);
                M += S->pending[g]->spec.m_raw;
            }
        }
        PrepGroup *grp = new PrepGroup();
        for (uint32_t g = 0; g < G; ++g) {
            Req *q = S->pending.front();
            S->pending.pop_front();
            const uint32_t nq = request_rows(q->spec);
            const ResLayout L = res_layout(nq, q->spec.m_raw);
            char *blk = res[g];
            q->resident = blk;
            q->off = (uint64_t *)blk;
            q->col = (uint32_t *)(blk + L.col);
            q->in_off = (uint64_t *)(blk + L.in_off);
            q->in_col = (uint32_t *)(blk + L.in_col);
            q->col_seg = (uint64_t *)(blk + L.col_seg);
            q->in_seg = (uint64_t *)(blk + L.in_seg);
            q->n_rows = nq;
            q->m_cap = q->spec.m_raw;
            grp->reqs.push_back(q);
        }
        grp->transient = t;
        grp->flag_slot = S->free_flags.back();
        S->free_flags.pop_back();
        S->preparing.push_back(grp);  // from here on fail_all releases its requests
        hipStream_t st = S->prep_stream[S->next_prep];
        S->next_prep = (S->next_prep + 1) % kPrepStreams;
        uint32_t *src = (uint32_t *)t, *dst = (uint32_t *)(t + 4 * M);
        char *wsp = t + round256(8 * M);
        uint32_t id0 = 0;
        uint64_t e0 = 0;
        for (Req *q : grp->reqs) {  // graph g generates ids [id0, id0 + n) into its slice of the union
            if (q->spec.kind == LSTC_REQ_RMAT)
                LSTC_TRY(gen_rmat_launch(q->spec.scale, q->spec.m_raw, q->spec.a, q->spec.b, q->spec.c, q->spec.seed,
                                         id0, src + e0, dst + e0, st));
            else
                LSTC_TRY(gen_er_launch(q->n_rows, q->spec.m_raw, q->spec.seed, id0, src + e0, dst + e0, st));
            id0 += q->n_rows;
            e0 += q->spec.m_raw;
        }
        // compact = 0 keeps every id, so the count needs no host-side sizes (and in a
        // union every graph keeps its id range)
        if (G == 1) {
            Req *q = grp->reqs[0];
            LSTC_TRY(clean_orient_launch(src, dst, M, (uint32_t)N, 0, nullptr, nullptr, q->off, q->col, q->in_off,
                                         q->in_col, q->col_seg, q->in_seg, wsp, ws, st, nullptr, &grp->d_bad));
        } else {
            char *u = wsp + round256(ws);
            const ResLayout UL = res_layout((uint32_t)N, M);
            uint64_t *uoff = (uint64_t *)u, *uin_off = (uint64_t *)(u + UL.in_off);
            uint32_t *ucol = (uint32_t *)(u + UL.col), *uin_col = (uint32_t *)(u + UL.in_col);
            uint64_t *useg = (uint64_t *)(u + UL.col_seg), *uin_seg = (uint64_t *)(u + UL.in_seg);
            LSTC_TRY(clean_orient_launch(src, dst, M, (uint32_t)N, 0, nullptr, nullptr, uoff, ucol, uin_off, uin_col,
                                         useg, uin_seg, wsp, ws, st, nullptr, &grp->d_bad));
            GroupTable T{};
            T.G = G;
            uint32_t b = 0;
            for (uint32_t g = 0; g < G; ++g) {
                T.base[g] = b;
                T.res[g] = (uint64_t)(grp->reqs[g]->resident - S->slab);
                T.mcap[g] = grp->reqs[g]->m_cap;
                b += grp->reqs[g]->n_rows;
            }
            T.base[G] = b;
            const uint32_t cus = (uint32_t)device_cus();
            const uint32_t grid_r = (uint32_t)hmin<uint64_t>((N + 255) / 256, (uint64_t)cus * 8);
            const uint32_t grid_a = (uint32_t)hmin<uint64_t>((M + 255) / 256, (uint64_t)cus * 16);
            hipLaunchKernelGGL(k_group_rows, dim3(grid_r), dim3(256), 0, st, T, S->slab, (const uint64_t *)uoff,
                               (const uint64_t *)uin_off);
            LSTC_LAUNCHED("k_group_rows");
            hipLaunchKernelGGL(k_group_arcs, dim3(grid_a), dim3(256), 0, st, T, S->slab, (const uint64_t *)uoff,
                               (const uint32_t *)ucol, (const uint64_t *)useg, (const uint64_t *)uin_off,
                               (const uint32_t *)uin_col, (const uint64_t *)uin_seg);
            LSTC_LAUNCHED("k_group_arcs");
        }
        // the out-of-range flag of the cleaning comes back with the preparation
        LSTC_HIP(hipMemcpyAsync(S->h_flags + grp->flag_slot, grp->d_bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        LSTC_HIP(hipEventCreateWithFlags(&grp->done, hipEventDisableTiming));
        LSTC_HIP(hipEventRecord(grp->done, st));
    }
    return LSTC_OK;
}

static void publish_failed(lstc_server *S, Req *r, lstc_status st) {
    const double lat = std::chrono::duration<double>(Clock::now() - r->t_submit).count();
    std::lock_guard<std::mutex> lk(S->mu);
    S->results[r->id] = Result{st, 0, 0, lat};
    S->n_completed++;
}

// Groups on different prep streams finish in any order: every one whose event
// has completed is collected.
static lstc_status collect_prepared(lstc_server *S) {
    bool failed_any = false;
    for (size_t i = 0; i < S->preparing.size();) {
        PrepGroup *grp = S->preparing[i];
        hipError_t e = hipEventQuery(grp->done);
        if (e == hipErrorNotReady) { ++i; continue; }
        if (e != hipSuccess) return hip_fail(e, "server: prep event");
        (void)hipEventDestroy(grp->done);
        S->pool.release(grp->transient);
        const uint32_t bad = S->h_flags[grp->flag_slot];
        S->free_flags.push_back(grp->flag_slot);
        S->preparing[i] = S->preparing.back();
        S->preparing.pop_back();
        for (Req *r : grp->reqs) {
            if (bad) {  // an id out of range (or a failed in-list scan): the graphs are not counted
                S->pool.release(r->resident);
                publish_failed(S, r, (bad & 2u) ? LSTC_ERR_DEVICE : LSTC_ERR_INVALID);
                delete r;
                failed_any = true;
            } else {
                S->ready.push_back(r);
            }
        }
        delete grp;
    }
    if (failed_any) S->cv_done.notify_all();
    return LSTC_OK;
}

size_t batch_head_bytes(uint32_t G) {
    Carver c(nullptr, 0);
    c.take<GraphDev>(G ? G : 1);
    c.take<uint64_t>((size_t)G + 1);
    c.take<unsigned long long>(G ? G : 1);
    c.take<unsigned long long>(G ? G : 1);
    return c.need();
}

static lstc_status launch_batch(lstc_server *S) {
    BatchSlot &B = S->slot[S->next_slot];
    if (!B.batch.empty() || S->ready.empty()) return LSTC_OK;
    // every ready request joins, as long as the work queue fits the slot's reserve
    std::vector<Req *> batch;
    uint64_t rows = 0, arcs = 0;
    while (!S->ready.empty() && batch.size() < S->max_batch) {
        Req *r = S->ready.front();
        const uint32_t G1 = (uint32_t)batch.size() + 1;
        if (!batch.empty() &&
            batch_head_bytes(G1) + count_ws_bytes(G1, rows + r->n_rows, arcs + r->m_cap) > S->reserve_bytes)
            break;
        batch.push_back(r);
        rows += r->n_rows;
        arcs += r->m_cap;
        S->ready.pop_front();
    }
    const uint32_t G = (uint32_t)batch.size();
    std::vector<HostGraph> hosts(G);
    uint64_t acc = 0;
    for (uint32_t g = 0; g < G; ++g) {
        Req *r = batch[g];
        hosts[g] = HostGraph{r->off, r->col, nullptr, r->n_rows, r->n_rows, r->m_cap, r->in_off, r->in_col,
                             r->col_seg, r->in_seg};
        B.h_desc[g] = to_dev(hosts[g]);
        B.h_base[g] = acc;
        acc += r->n_rows;
    }
    B.h_base[G] = acc;
    const size_t head = batch_head_bytes(G);
    Carver d(B.reserve, S->reserve_bytes);
    GraphDev *d_desc = d.take<GraphDev>(G);
    uint64_t *d_base = d.take<uint64_t>((size_t)G + 1);
    unsigned long long *d_counts = d.take<unsigned long long>(G);
    unsigned long long *d_edges = d.take<unsigned long long>(G);
    hipStream_t st = S->count_stream;
    B.batch = batch;  // from here on the slot owns the requests (fail_all releases them)
    LSTC_HIP(hipMemcpyAsync(d_desc, B.h_desc, sizeof(GraphDev) * G, hipMemcpyHostToDevice, st));
    LSTC_HIP(hipMemcpyAsync(d_base, B.h_base, sizeof(uint64_t) * (G + 1), hipMemcpyHostToDevice, st));
    LSTC_TRY(count_launch(hosts.data(), G, G > 1 ? d_desc : nullptr, G > 1 ? d_base : nullptr, d_counts,
                          S->has_tune ? &S->tune : nullptr, B.reserve + head, S->reserve_bytes - head, st));
    hipLaunchKernelGGL(k_gather_edges, dim3(1), dim3(256), 0, st, d_desc, G, d_edges);
    LSTC_LAUNCHED("k_gather_edges");
    LSTC_HIP(hipMemcpyAsync(B.h_counts, d_counts, sizeof(unsigned long long) * G, hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipMemcpyAsync(B.h_edges, d_edges, sizeof(unsigned long long) * G, hipMemcpyDeviceToHost, st));
    LSTC_HIP(hipEventRecord(B.done, st));
    S->next_slot = (S->next_slot + 1) % kSlots;
    return LSTC_OK;
}

// Slots complete in launch order (one count stream): the oldest one is checked.
static lstc_status finish_batch(lstc_server *S) {
    for (int k = 0; k < kSlots; ++k) {
        BatchSlot &B = S->slot[(S->next_slot + k) % kSlots];
        if (B.batch.empty()) continue;
        hipError_t e = hipEventQuery(B.done);
        if (e == hipErrorNotReady) return LSTC_OK;
        if (e != hipSuccess) return hip_fail(e, "server: batch event");
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (size_t g = 0; g < B.batch.size(); ++g) {
                Req *r = B.batch[g];
                const double lat = std::chrono::duration<double>(now - r->t_submit).count();
                S->results[r->id] = Result{LSTC_OK, (uint64_t)B.h_counts[g], (uint64_t)B.h_edges[g], lat};
                S->latencies.push_back(lat);
                S->n_completed++;
                S->edges_done += B.h_edges[g];
                S->triangles_done += B.h_counts[g];
            }
            S->n_batches++;
            S->max_batch_seen = std::max<uint64_t>(S->max_batch_seen, B.batch.size());
        }
        for (Req *r : B.batch) {
            S->pool.release(r->resident);
            delete r;
        }
        B.batch.clear();
        S->cv_done.notify_all();
    }
    return LSTC_OK;
}

static void fail_all(lstc_server *S, lstc_status st) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->fatal = st;
    auto drop = [&](Req *r) {
        S->results[r->id] = Result{st, 0, 0, 0.0};
        delete r;
    };
    for (Req *r : S->pending) drop(r);
    for (PrepGroup *g : S->preparing) {
        for (Req *r : g->reqs) drop(r);
        delete g;
    }
    for (Req *r : S->ready) drop(r);
    for (BatchSlot &B : S->slot) {
        for (Req *r : B.batch) drop(r);
        B.batch.clear();
    }
    for (Req *r : S->submitted) drop(r);
    S->pending.clear(); S->preparing.clear(); S->ready.clear(); S->submitted.clear();
    S->queued.clear(); S->cancelled.clear();
    S->cv_done.notify_all();
}

static bool slots_busy(const lstc_server *S) {
    for (const BatchSlot &B : S->slot)
        if (!B.batch.empty()) return true;
    return false;
}

static void worker_main(lstc_server *S) {
    if (hipSetDevice(S->device) != hipSuccess) { fail_all(S, LSTC_ERR_HIP); return; }
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            const bool idle = S->pending.empty() && S->preparing.empty() && S->ready.empty() && !slots_busy(S);
            if (idle && S->submitted.empty()) {
                if (S->stop) break;
                S->cv_work.wait_for(lk, std::chrono::milliseconds(2));
            }
            while (!S->submitted.empty()) {
                S->pending.push_back(S->submitted.front());
                S->submitted.pop_front();
            }
        }
        lstc_status rc = finish_batch(S);
        if (rc == LSTC_OK) rc = collect_prepared(S);
        if (rc == LSTC_OK) rc = admit(S);
        for (int k = 0; rc == LSTC_OK && k < kSlots; ++k) rc = launch_batch(S);
        if (rc != LSTC_OK) { fail_all(S, rc); break; }
        const bool busy = slots_busy(S) || !S->preparing.empty();
        if (busy) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

}  // namespace lstc

extern "C" lstc_status lstc_server_create(size_t hbm_budget_bytes, uint32_t max_batch, const lstc_tune *tune,
                                          lstc_server **out) {
    if (!out) return fail(LSTC_ERR_INVALID, "lstc_server_create: NULL out");
    *out = nullptr;
    if (hbm_budget_bytes < (1u << 20)) return fail(LSTC_ERR_INVALID, "lstc_server_create: budget below 1 MiB");
    if (max_batch == 0) max_batch = 1024;
    lstc_server *S = new lstc_server();
    S->budget = hbm_budget_bytes;
    S->max_batch = max_batch;
    if (tune) { S->tune = *tune; S->has_tune = true; }
    hipError_t e = hipGetDevice(&S->device);
    if (e == hipSuccess) e = hipMalloc((void **)&S->slab, hbm_budget_bytes);
    if (e != hipSuccess) {
        delete S;
        return fail(LSTC_ERR_OOM, "lstc_server_create: cannot allocate %zu bytes of HBM", hbm_budget_bytes);
    }
    // batch work queues live in a fixed reserve (one share per batch slot), so resident
    // CSRs can never starve the count
    size_t reserve = hmin<size_t>(hmax<size_t>(hbm_budget_bytes / 8, (size_t)32 << 20), (size_t)4 << 30);
    reserve = hmin<size_t>(reserve, hbm_budget_bytes / 2);
    S->reserve_bytes = (reserve / kSlots) & ~size_t(255);
    char *res0 = S->slab + (hbm_budget_bytes - (size_t)kSlots * S->reserve_bytes);
    S->pool.init(S->slab, (size_t)(res0 - S->slab) & ~size_t(255));
    for (int i = 0; i < kPrepStreams && e == hipSuccess; ++i)
        e = hipStreamCreateWithFlags(&S->prep_stream[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->count_stream, hipStreamNonBlocking);
    for (int k = 0; k < kSlots && e == hipSuccess; ++k) {
        BatchSlot &B = S->slot[k];
        B.reserve = res0 + (size_t)k * S->reserve_bytes;
        e = hipHostMalloc((void **)&B.h_desc, sizeof(GraphDev) * max_batch);
        if (e == hipSuccess) e = hipHostMalloc((void **)&B.h_base, sizeof(uint64_t) * (max_batch + 1));
        if (e == hipSuccess) e = hipHostMalloc((void **)&B.h_counts, sizeof(unsigned long long) * max_batch);
        if (e == hipSuccess) e = hipHostMalloc((void **)&B.h_edges, sizeof(unsigned long long) * max_batch);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&B.done, hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipHostMalloc((void **)&S->h_flags, sizeof(uint32_t) * kFlagSlots);
    if (e != hipSuccess) {
        lstc_server_destroy(S);
        return hip_fail(e, "lstc_server_create");
    }
    for (int i = kFlagSlots - 1; i >= 0; --i) S->free_flags.push_back(i);
    S->worker = std::thread(worker_main, S);
    *out = S;
    return LSTC_OK;
}

extern "C" lstc_status lstc_server_submit(lstc_server *S, const lstc_request *req, uint64_t *id) {
    if (!S || !req || !id) return fail(LSTC_ERR_INVALID, "lstc_server_submit: NULL argument");
    const lstc_request &r = *req;
    if (r.kind != LSTC_REQ_RMAT && r.kind != LSTC_REQ_ER && r.kind != LSTC_REQ_CSR)
        return fail(LSTC_ERR_INVALID, "lstc_server_submit: unknown request kind %u", r.kind);
    if (r.kind == LSTC_REQ_RMAT && (r.scale == 0 || r.scale > 31))
        return fail(LSTC_ERR_INVALID, "lstc_server_submit: RMAT scale %u not in 1..31", r.scale);
    if (r.kind == LSTC_REQ_RMAT && !(r.a >= 0 && r.b >= 0 && r.c >= 0 && r.a + r.b + r.c <= 1.0))
        return fail(LSTC_ERR_INVALID, "lstc_server_submit: RMAT probabilities a, b, c must be >= 0 with a+b+c <= 1");
    if (r.kind == LSTC_REQ_ER && (r.n == 0 || r.n == 0xFFFFFFFFu))
        return fail(LSTC_ERR_INVALID, "lstc_server_submit: bad ER vertex count");
    if (r.kind == LSTC_REQ_CSR) {
        HostGraph h{r.off, r.col, nullptr, r.n_rows, r.n_rows, r.m};
        LSTC_TRY(validate_graph(h, "lstc_server_submit"));
        if (batch_head_bytes(1) + count_ws_bytes(1, r.n_rows, r.m) > S->reserve_bytes)
            return fail(LSTC_ERR_OVER_BUDGET, "lstc_server_submit: CSR rows exceed the batch reserve");
    } else {
        if (r.m_raw >= 0x7FFFF000ull) return fail(LSTC_ERR_INVALID, "lstc_server_submit: too many edges");
        // the working set of one request must fit the budget on its own
        const uint32_t n = request_rows(r);
        size_t ws = 0;
        const size_t need = prep_bytes(n, r.m_raw, &ws) + resident_bytes(n, r.m_raw);
        const size_t qneed = batch_head_bytes(1) + count_ws_bytes(1, n, r.m_raw);
        if (need > S->pool.total() || qneed > S->reserve_bytes) {
            std::lock_guard<std::mutex> lk(S->mu);
            S->n_rejected++;
            return fail(LSTC_ERR_OVER_BUDGET, "lstc_server_submit: request needs %zu + %zu bytes, budget is %zu",
                        need, qneed, S->budget);
        }
    }
    Req *q = new Req();
    q->spec = r;
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->fatal != LSTC_OK) { delete q; return fail(S->fatal, "lstc_server_submit: server failed"); }
        q->id = S->next_id++;
        S->n_submitted++;
        S->submitted.push_back(q);
        S->queued.insert(q->id);
        *id = q->id;
    }
    S->cv_work.notify_one();
    return LSTC_OK;
}

static lstc_status take_result(lstc_server *S, std::unordered_map<uint64_t, Result>::iterator it, lstc_result *out) {
    if (out) {
        out->count = it->second.count;
        out->edges = it->second.edges;
        out->latency_s = it->second.latency_s;
    }
    const lstc_status st = it->second.status;
    S->results.erase(it);
    return st;
}

extern "C" lstc_status lstc_server_poll(lstc_server *S, uint64_t id, lstc_result *out) {
    if (!S) return fail(LSTC_ERR_INVALID, "lstc_server_poll: NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->results.find(id);
    if (it == S->results.end()) return LSTC_ERR_NOT_READY;
    return take_result(S, it, out);
}

extern "C" lstc_status lstc_server_wait(lstc_server *S, uint64_t id, double timeout_s, lstc_result *out) {
    if (!S) return fail(LSTC_ERR_INVALID, "lstc_server_wait: NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    for (;;) {
        auto it = S->results.find(id);
        if (it != S->results.end()) return take_result(S, it, out);
        if (S->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && Clock::now() >= deadline)
            return fail(LSTC_ERR_NOT_READY, "lstc_server_wait: timed out");
    }
}

extern "C" lstc_status lstc_server_cancel(lstc_server *S, uint64_t id) {
    if (!S) return fail(LSTC_ERR_INVALID, "lstc_server_cancel: NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (id == 0 || id >= S->next_id) return fail(LSTC_ERR_INVALID, "lstc_server_cancel: unknown request %llu",
                                                 (unsigned long long)id);
    if (!S->queued.count(id))
        return fail(LSTC_ERR_NOT_READY, "lstc_server_cancel: request %llu is already admitted or finished",
                    (unsigned long long)id);
    S->cancelled.insert(id);
    S->cv_work.notify_one();
    return LSTC_OK;
}

extern "C" lstc_status lstc_server_get_stats(lstc_server *S, lstc_server_stats *out) {
    if (!S || !out) return fail(LSTC_ERR_INVALID, "lstc_server_get_stats: NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out->submitted = S->n_submitted;
    out->completed = S->n_completed;
    out->rejected = S->n_rejected;
    out->batches = S->n_batches;
    out->max_batch_graphs = S->max_batch_seen;
    out->pool_bytes = S->pool.total();
    out->pool_used_bytes = S->pool.used();
    out->edges = S->edges_done;
    out->triangles = S->triangles_done;
    std::vector<double> l = S->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
        if (l.empty()) return 0.0;
        size_t i = (size_t)(p * (double)(l.size() - 1) + 0.5);
        return l[std::min(i, l.size() - 1)];
    };
    out->p50_s = pct(0.50);
    out->p99_s = pct(0.99);
    out->max_s = l.empty() ? 0.0 : l.back();
    double s = 0;
    for (double x : l) s += x;
    out->mean_s = l.empty() ? 0.0 : s / (double)l.size();
    out->prep_groups = S->n_groups;
    out->max_prep_graphs = S->max_group_seen;
    return LSTC_OK;
}

extern "C" lstc_status lstc_server_destroy(lstc_server *S) {
    if (!S) return LSTC_OK;
    if (S->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
        }
        S->cv_work.notify_all();
        S->worker.join();
    }
    for (hipStream_t &ps : S->prep_stream)
        if (ps) { (void)hipStreamSynchronize(ps); (void)hipStreamDestroy(ps); }
    if (S->count_stream) { (void)hipStreamSynchronize(S->count_stream); (void)hipStreamDestroy(S->count_stream); }
    for (BatchSlot &B : S->slot) {
        if (B.done) (void)hipEventDestroy(B.done);
        if (B.h_desc) (void)hipHostFree(B.h_desc);
        if (B.h_base) (void)hipHostFree(B.h_base);
        if (B.h_counts) (void)hipHostFree(B.h_counts);
        if (B.h_edges) (void)hipHostFree(B.h_edges);
    }
    if (S->h_flags) (void)hipHostFree(S->h_flags);
    if (S->slab) (void)hipFree(S->slab);
    delete S;
    return LSTC_OK;
}
