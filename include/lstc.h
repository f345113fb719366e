/*
 * lstc.h -- C ABI of liblstc, the MI355X (gfx950) exact triangle-counting path.
 *
 * Conventions
 *   - Plain C types only; device pointers are ordinary pointers into HBM,
 *     streams are hipStream_t passed as void* (NULL = default stream).
 *   - The caller owns every input and output buffer.  Calls that need scratch
 *     take a caller-provided workspace sized by the matching *_workspace call
 *     (the two-phase temp-storage idiom of the reference's CUB calls,
 *     src/WithoutPartition/ShareKernelTCV62_update_v1.cu:271-277).
 *   - Every call returns lstc_status (0 = ok, negative = error) and never exits
 *     or prints.  lstc_last_error() gives a detail string for this thread.
 *   - Calls are stream-ordered; those that must report host-side sizes
 *     (lstc_clean_orient, lstc_halo_sizes) synchronise their stream first.
 *   - Vertex ids are uint32 (< 2^32 - 1), CSR row offsets uint64, counts uint64.
 *
 * The reference exposes no library API; its boundary is executables plus text
 * files.  Each entry point below names the reference code it replaces.
 */
#ifndef LSTC_H
#define LSTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lstc_status;
#define LSTC_OK 0
#define LSTC_ERR_INVALID (-1)     /* bad argument / shape */
#define LSTC_ERR_OOM (-2)         /* device allocation failed (server only) */
#define LSTC_ERR_OVER_BUDGET (-3) /* request larger than the HBM budget */
#define LSTC_ERR_HIP (-4)         /* HIP runtime error (see lstc_last_error) */
#define LSTC_ERR_WORKSPACE (-5)   /* workspace missing or too small */
#define LSTC_ERR_NOT_READY (-6)   /* result not yet available (server poll) */
#define LSTC_ERR_CANCELLED (-7)   /* request cancelled or timed out */

typedef void *lstc_stream; /* hipStream_t */

const char *lstc_status_string(lstc_status st);
const char *lstc_last_error(void);
int lstc_version(void); /* major*10000 + minor*100 + patch */

/* ------------------------------------------------------------------ generators
 * Seeded, counter-based (SplitMix64) synthetic inputs written straight into
 * HBM.  The reference reads public datasets from disk instead (README.md:14-18).
 */
lstc_status lstc_gen_rmat(uint32_t scale, uint64_t m_raw, double a, double b, double c,
                          uint64_t seed, uint32_t *d_src, uint32_t *d_dst, lstc_stream s);
lstc_status lstc_gen_er(uint32_t n, uint64_t m_raw, uint64_t seed, uint32_t *d_src,
                        uint32_t *d_dst, lstc_stream s);
lstc_status lstc_gen_chung_lu_workspace(uint32_t n, size_t *bytes);
lstc_status lstc_gen_chung_lu(uint32_t n, double gamma, double mean_deg, double w_max,
                              uint64_t m_raw, uint64_t seed, uint32_t *d_src, uint32_t *d_dst,
                              void *d_ws, size_t ws_bytes, lstc_stream s);
/* rows x cols grid, 3 raw slots per vertex (right, down, diagonal with prob p);
 * missing slots are self loops that cleaning removes.  m_raw = 3*rows*cols.   */
lstc_status lstc_gen_grid(uint32_t rows, uint32_t cols, double p_diag, uint64_t seed,
                          uint32_t *d_src, uint32_t *d_dst, lstc_stream s);

/* Seeded uniform k-way vertex assignment; replaces the unseeded
 * np.random.randint of preprocess_GPU/generate_clean_undir_dataset/
 * Halo_creation_preprocess.py:288,325,362.                                   */
lstc_status lstc_partition_random(uint32_t n, uint32_t k, uint64_t seed, uint32_t *d_part,
                                  lstc_stream s);

/* ------------------------------------------------------------------ (a) clean + orient
 * Replaces Med_dataset/without_part_preprocess.py:193-208 (DGL remove_self_loop,
 * to_bidirected, isolated-node removal) and :19-111,253-270 (Find_size, cp.cumsum,
 * Convert).  Drops self loops and repeats, symmetrises, optionally (compact=1)
 * removes isolated vertices with an order-preserving relabel, and orients every
 * edge s->d iff deg(s) < deg(d) or (deg equal and d < s).
 * Outputs: optional symmetric CSR (d_sym_off[n+1], d_sym_col[2*m_raw]) and the
 * oriented CSR (d_off[n+1], d_col[m_raw]); *n_out / *m_out receive the vertex
 * and undirected edge counts (the stream is synchronised).                    */
lstc_status lstc_clean_orient_workspace(uint32_t n, uint64_t m_raw, size_t *bytes);
lstc_status lstc_clean_orient(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                              uint32_t n, int compact, uint64_t *d_sym_off, uint32_t *d_sym_col,
                              uint64_t *d_off, uint32_t *d_col, uint32_t *n_out, uint64_t *m_out,
                              void *d_ws, size_t ws_bytes, lstc_stream s);

/* ------------------------------------------------------------------ (c) counting
 * Tuning knobs (picked by lstc_autotune; zero-initialised = defaults).       */
typedef struct lstc_tune {
    uint32_t wg_waves;      /* waves per workgroup (1..16), 0 = 4                */
    uint32_t table_scale;   /* LDS hash capacity = table_scale * keys (2..8), 0 = 4 */
    uint32_t switch_ratio;  /* per-arc side switch: iterate N+(u) against a hashed
                               N+(w) when d+(u)*switch_ratio/16 < d+(w); 0 = off  */
    uint32_t split_segments;/* max neighbour segments per work item (multiple of 64), 0 = 256 */
    uint32_t reserved[4];
} lstc_tune;

/* Exact count over owned rows [0, n_owned) of an oriented CSR: the sum over
 * owned u of sum_{w in N+(u)} |N+(u) n N+(w)|.  Replaces Find_Triangle +
 * cub::DeviceReduce::Sum (src/WithoutPartition/ShareKernelTCV62_update_v1.cu:43-191,
 * 264, 271-277) and, with n_owned < n_rows, the owned-row launch of
 * src/WithPartition/without_sorting_tc.cu:57-219, 400, 422-428.
 * m = d_off[n_rows] is the arc count (the reference kernel's `edge` argument),
 * limited to m < 2^32 - 4096 by the 32-bit work positions inside a row.
 * d_count (1 uint64) is overwritten.  d_per_row (n_owned uint64, may be NULL)
 * receives the reference's per-row partials g_sum[u].                         */
lstc_status lstc_count_workspace(uint32_t n_rows, uint32_t n_owned, uint64_t m,
                                 const lstc_tune *tune, size_t *bytes);
lstc_status lstc_count(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows,
                       uint32_t n_owned, uint64_t m, uint64_t *d_count, uint64_t *d_per_row,
                       const lstc_tune *tune, void *d_ws, size_t ws_bytes, lstc_stream s);

#ifdef __cplusplus
}
#endif
#endif /* LSTC_H */
