/*
 * lstc_oracle.h -- CPU restatement of the LSTC triangle-counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (trianglecount_amd/,
 * include/) links, loads or calls this code.  It checks results in tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Paths are relative to the reference repository root.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   lo_gen_*            : closed-form structure (grid), determinism, and
 *                         statistical shape; bit-equality vs the HIP generator.
 *   lo_clean            : brute-force std-set semantics, networkx on small
 *                         graphs, golden toy CSRs from test_code/.
 *   lo_orient           : the compiled reference preprocess_CPU/Undir2Dir.c
 *                         (oracle/_ref), golden vectors, DAG/rank invariants.
 *   lo_count_merge      : closed forms (K_n, wheel, windmill, grid 2*D),
 *                         networkx triangles, scipy trace(A^3)/6, brute force.
 *   lo_count_rows_k1    : equality with lo_count_merge per row on random graphs
 *                         (two different algorithms), closed forms.
 *   lo_halo_extract     : brute-force definition on small graphs, and the
 *                         invariant sum over parts == unpartitioned count.
 *   lo_triangle_support : support of edge {u,w} = |N(u) n N(w)| over the symmetric
 *                         graph, per-vertex counts = networkx triangles, closed forms
 *                         (K_n, wheel), sums = 3 x lo_count_merge.
 *   lo_partition_random : determinism, range, balance.
 */
#ifndef LSTC_ORACLE_H
#define LSTC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG (SplitMix64 finaliser, Steele et al. 2014) ---- */
uint64_t lo_mix64(uint64_t z);
uint64_t lo_stream_key(uint64_t seed, uint64_t stream);
uint64_t lo_rng64(uint64_t key, uint64_t ctr);

/* ---- synthetic generators (raw, uncleaned edge samples) ---- */
void lo_gen_rmat(uint32_t scale, uint64_t m_raw, double a, double b, double c,
                 uint64_t seed, uint32_t *src, uint32_t *dst);
void lo_gen_er(uint32_t n, uint64_t m_raw, uint64_t seed, uint32_t *src, uint32_t *dst);
/* Chung-Lu: returns 0 on success; writes the weight table total to *total_w (may be NULL). */
int lo_gen_chung_lu(uint32_t n, double gamma, double mean_deg, double w_max, uint64_t m_raw,
                    uint64_t seed, uint32_t *src, uint32_t *dst);
/* 2-D grid, 3 slots per vertex (right, down, diagonal); absent -> self loop. m_raw = 3*R*C. */
void lo_gen_grid(uint32_t rows, uint32_t cols, double p_diag, uint64_t seed,
                 uint32_t *src, uint32_t *dst);
uint64_t lo_grid_num_diagonals(uint32_t rows, uint32_t cols, double p_diag, uint64_t seed);

/* ---- (a) clean: drop loops, symmetrise, dedupe, optionally drop isolated ----
 * Output: symmetric CSR with sorted rows.  sym_off has n_out+1 entries and
 * sym_col has 2*m entries (caller allocates n+1 and 2*m_raw).  Returns n_out
 * through *n_out and the undirected edge count through *m_out.            */
void lo_clean(uint32_t n, uint64_t m_raw, const uint32_t *src, const uint32_t *dst,
              int compact, uint32_t *n_out, uint64_t *m_out,
              uint64_t *sym_off, uint32_t *sym_col);

/* ---- (a) orient by (degree, id): keep s->d iff deg s < deg d or (equal and d < s) ----
 * Input symmetric CSR, output oriented CSR (off n+1, col m).              */
void lo_orient(uint32_t n, const uint64_t *sym_off, const uint32_t *sym_col,
               uint64_t *off, uint32_t *col);

/* ---- (c) counting ---- */
/* Merge-intersection counter over owned rows [0, n_owned) of an oriented CSR.
 * per_row (may be NULL) receives |N+(u) n N+(w)| summed over w in N+(u).
 * threads <= 0 means use all OpenMP threads.                              */
uint64_t lo_count_merge(uint32_t n_rows, uint32_t n_owned, const uint64_t *off,
                        const uint32_t *col, uint64_t *per_row, int threads);
/* The same counter over an explicit row list (used for bounded CPU samples). */
uint64_t lo_count_merge_list(const uint64_t *off, const uint32_t *col, const uint32_t *rows,
                             uint64_t nrows, int threads);
/* Literal restatement of the reference block-per-row kernel semantics
 * (tiles of 128 staged entries, endpoint-then-bisection search).          */
uint64_t lo_count_rows_k1(uint32_t n_rows, uint32_t n_owned, const uint64_t *off,
                          const uint32_t *col, uint64_t *g_sum);

/* Per-edge support (support[m], by arc position) and per-vertex triangle counts
 * (per_vertex[n_rows]) over the triangles whose source row is owned; either
 * output may be NULL.                                                      */
void lo_triangle_support(uint32_t n_rows, uint32_t n_owned, const uint64_t *off, const uint32_t *col,
                         uint32_t *support, uint64_t *per_vertex);

/* ---- partition and (b) halo extraction ---- */
void lo_partition_random(uint32_t n, uint32_t k, uint64_t seed, uint32_t *part);
/* mode 0: vertex set P u N(P) (the reference's un-pruned output);
 * mode 1: P u N+(P) (count-preserving pruning).
 * Sizes pass: call with off_out == NULL to get *n_local, *n_owned, *m_local.
 * Fill pass: off_out (n_local+1), col_out (m_local), l2g (n_local, may be NULL). */
void lo_halo_extract(uint32_t n, const uint64_t *sym_off, const uint32_t *sym_col,
                     const uint32_t *part, uint32_t p, int mode,
                     uint32_t *n_local, uint32_t *n_owned, uint64_t *m_local,
                     uint64_t *off_out, uint32_t *col_out, uint32_t *l2g);

/* ---- statistics for the roofline model ---- */
typedef struct {
    uint64_t n, m;
    uint64_t max_out;        /* max d+(u) */
    uint64_t sum_out_sq;     /* sum_u d+(u)^2 */
    uint64_t sum_in_out;     /* sum_v d-(v) d+(v) */
    uint64_t rows_ge2;       /* rows with d+ >= 2 */
} lo_stats_t;
void lo_stats(uint32_t n, const uint64_t *off, const uint32_t *col, lo_stats_t *st);

#ifdef __cplusplus
}
#endif
#endif
