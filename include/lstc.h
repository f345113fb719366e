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
#define LSTC_ERR_DEVICE (-8)      /* a device-side wait exceeded its bound; the result is not valid */

typedef void *lstc_stream; /* hipStream_t */

const char *lstc_status_string(lstc_status st);
const char *lstc_last_error(void);
int lstc_version(void); /* major*10000 + minor*100 + patch */

/* Device self-test of the wave64 primitives (DPP prefix sum, reduction, lane
 * permute, ballot rank) that the counting kernel's addressing is built on: runs
 * them on fixed patterns on the current device and compares with host
 * arithmetic.  *mismatches receives the number of wrong lanes (0 = pass).
 * Diagnostic only; it has no counterpart in the reference.                    */
lstc_status lstc_selftest(uint32_t *mismatches);

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
/* Contiguous blocks, Halo_creation_preprocess.py:292-300: l = n / k, part i is
 * [i*l, (i+1)*l), the remainder (everything when n < k) stays in part 0.     */
lstc_status lstc_partition_contig(uint32_t n, uint32_t k, uint32_t *d_part, lstc_stream s);
/* An externally computed assignment (e.g. METIS) from a node-parts text file:
 * one integer per vertex, the format Halo_creation_preprocess.py:308-312
 * writes and round_clean_subgraph.cu:964-968 reads.  The first n values are
 * uploaded into d_part[n]; a short file or a value >= k is refused with
 * LSTC_ERR_INVALID.  Synchronises the stream.                                */
lstc_status lstc_partition_read(const char *path, uint32_t n, uint32_t k, uint32_t *d_part, lstc_stream s);

/* ------------------------------------------------------------------ (a) clean + orient
 * Replaces Med_dataset/without_part_preprocess.py:193-208 (DGL remove_self_loop,
 * to_bidirected, isolated-node removal) and :19-111,253-270 (Find_size, cp.cumsum,
 * Convert).  Drops self loops and repeats, symmetrises, optionally (compact=1)
 * removes isolated vertices with an order-preserving relabel, and orients every
 * edge s->d iff deg(s) < deg(d) or (deg equal and d < s).
 * Outputs: optional symmetric CSR (d_sym_off[n+1], d_sym_col[2*m_raw]), the
 * oriented CSR (d_off[n+1], d_col[m_raw]) and optional "target-side in-lists"
 * (d_in_off[n+1], d_in_col[m_raw]): for every w, the sources u of arcs u->w with
 * d+(u) <= d+(w), ascending.  Those are the arcs lstc_count_split counts from
 * w's side.  *n_out / *m_out receive the vertex and undirected edge counts (the
 * stream is synchronised).                                                    */
lstc_status lstc_clean_orient_workspace(uint32_t n, uint64_t m_raw, size_t *bytes);
lstc_status lstc_clean_orient(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw,
                              uint32_t n, int compact, uint64_t *d_sym_off, uint32_t *d_sym_col,
                              uint64_t *d_off, uint32_t *d_col, uint64_t *d_in_off, uint32_t *d_in_col,
                              uint32_t *n_out, uint64_t *m_out, void *d_ws, size_t ws_bytes,
                              lstc_stream s);

/* lstc_clean_orient plus everything the split count reads (lstc_count_split):
 * the in-lists and the segment words of d_col and d_in_col (as lstc_segments
 * builds them), in one stream-ordered pass with one synchronisation at the end.
 * d_col_seg[m_raw], d_in_seg[m_raw]; *n_in_out receives the in-list entries.  */
lstc_status lstc_clean_orient_split(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw, uint32_t n,
                                    int compact, uint64_t *d_sym_off, uint32_t *d_sym_col, uint64_t *d_off,
                                    uint32_t *d_col, uint64_t *d_in_off, uint32_t *d_in_col, uint64_t *d_col_seg,
                                    uint64_t *d_in_seg, uint32_t *n_out, uint64_t *m_out, uint64_t *n_in_out,
                                    void *d_ws, size_t ws_bytes, lstc_stream s);

/* lstc_clean_orient_split plus a streaming copy of the oriented columns for
 * lstc_count_split_stream: d_scol holds every row again, each starting on a
 * 32-entry (128-byte) boundary (given a 128-byte aligned d_scol), and the
 * segment words d_col_seg / d_in_seg hold row starts in d_scol instead of d_col.
 * The count then reads each streamed list from its first cache line on (4.6%
 * faster at RMAT s23; same result).  d_scol needs lstc_stream_copy_capacity(n,
 * m_raw) = m_raw + 31 n entries (m_raw + 31 n < 2^32); *scol_out receives the
 * entries used.  No counterpart in the reference (it reads col directly).    */
uint64_t lstc_stream_copy_capacity(uint32_t n, uint64_t m_raw);
lstc_status lstc_clean_orient_stream(const uint32_t *d_src, const uint32_t *d_dst, uint64_t m_raw, uint32_t n,
                                     int compact, uint64_t *d_sym_off, uint32_t *d_sym_col, uint64_t *d_off,
                                     uint32_t *d_col, uint64_t *d_in_off, uint32_t *d_in_col, uint64_t *d_col_seg,
                                     uint64_t *d_in_seg, uint32_t *d_scol, uint32_t *n_out, uint64_t *m_out,
                                     uint64_t *n_in_out, uint64_t *scol_out, void *d_ws, size_t ws_bytes,
                                     lstc_stream s);

/* ------------------------------------------------------------------ (b) halo subgraphs
 * Replaces create_subgraph_with_halo (preprocess_GPU/Halo_cleanned_preprocess/
 * round_clean_subgraph.cu:357-902) for part p of a vertex assignment d_part[n]:
 * the induced subgraph on P u H, P = {v : part[v] == p} ascending, H = N(P)\P
 * (mode 0, the reference's un-pruned output) or N+(P)\P (mode 1, the
 * count-preserving pruning), local ids P block then H block, arcs oriented by the
 * GLOBAL (degree, id) rule, rows ascending.  Input: the symmetric CSR from
 * lstc_clean_orient.  Owned rows are the first n_owned = |P| rows, as in the
 * reference's file format (:850).  Both calls synchronise the stream.          */
lstc_status lstc_halo_workspace(uint32_t n, size_t *bytes);
lstc_status lstc_halo_sizes(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                            const uint32_t *d_part, uint32_t p, int mode, uint32_t *n_local,
                            uint32_t *n_owned, uint64_t *m_local, void *d_ws, size_t ws_bytes,
                            lstc_stream s);
/* d_off_out[n_local+1], d_col_out[m_local], d_l2g_out[n_local] (local -> global, may be NULL). */
lstc_status lstc_halo_extract(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                              const uint32_t *d_part, uint32_t p, int mode, uint64_t *d_off_out,
                              uint32_t *d_col_out, uint32_t *d_l2g_out, uint32_t n_local_cap,
                              uint64_t m_local_cap, void *d_ws, size_t ws_bytes, lstc_stream s);

/* The halo subgraphs of all k parts at once (1 <= k <= 64): part p's output is
 * exactly what lstc_halo_extract gives for p, but every pass over the
 * symmetric CSR serves all parts (a per-vertex mask of the parts whose local
 * set holds it), where the reference builds one part per pass
 * (round_clean_subgraph.cu:904-976, called once per part).  Vertices with
 * part[v] >= k are owned by no part.  Both calls synchronise the stream.     */
typedef struct lstc_halo_out {
    uint64_t *off;        /* device, n_local + 1                 */
    uint32_t *col;        /* device, m_local                     */
    uint32_t *l2g;        /* device, n_local (local -> global), or NULL for every part */
    uint32_t n_local_cap;
    uint32_t reserved;
    uint64_t m_local_cap;
} lstc_halo_out;
lstc_status lstc_halo_parts_workspace(uint32_t n, uint32_t k, size_t *bytes);
/* n_local[k], n_owned[k], m_local[k] */
lstc_status lstc_halo_parts_sizes(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                  const uint32_t *d_part, uint32_t k, int mode, uint32_t *n_local,
                                  uint32_t *n_owned, uint64_t *m_local, void *d_ws, size_t ws_bytes,
                                  lstc_stream s);
/* outs[k] (host array of device buffers with their capacities) */
lstc_status lstc_halo_parts_extract(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                    const uint32_t *d_part, uint32_t k, int mode, const lstc_halo_out *outs,
                                    void *d_ws, size_t ws_bytes, lstc_stream s);
/* Sizes and extraction in one pass over the graph, for callers that give each
 * part buffers of an upper bound (n_local <= n, m_local <= the graph's
 * undirected edges, sym_off[n] / 2): n_local[k], n_owned[k], m_local[k]
 * receive the sizes; if a part's capacity is too small nothing is written to
 * the outputs, the sizes are still returned, and the call fails with
 * LSTC_ERR_INVALID.                                                           */
lstc_status lstc_halo_parts_build(const uint64_t *d_sym_off, const uint32_t *d_sym_col, uint32_t n,
                                  const uint32_t *d_part, uint32_t k, int mode, const lstc_halo_out *outs,
                                  uint32_t *n_local, uint32_t *n_owned, uint64_t *m_local, void *d_ws,
                                  size_t ws_bytes, lstc_stream s);

/* ------------------------------------------------------------------ (c) counting
 * Tuning knobs (picked by lstc_autotune; zero-initialised = defaults).       */
typedef struct lstc_tune {
    uint32_t table_scale;  /* LDS hash capacity >= table_scale * keys (2..16), 0 = 6    */
    uint32_t small_max;    /* switch point: rows with d+ <= small_max get a wave-private
                              1024-slot table, larger rows a workgroup-shared 4096-slot
                              one (64..1024/table_scale); 0 = min(256, 1024/table_scale),
                              capped at 128 for a launch under 2^23 arcs              */
    uint32_t small_group;  /* small rows per work ticket (4..256), 0 = planner's choice */
    uint32_t grid_per_cu;  /* persistent workgroups per CU (1..32), 0 = 32 / waves     */
    uint32_t unroll;       /* 64-entry chunks in flight per wave (4 or 8), 0 = 4       */
    uint32_t tiny_max;     /* graphs with in-lists: rows with 2 <= d+ <= tiny_max are
                              counted one lane per arc, comparing the two lists (each
                              at most tiny_max <= 8 entries) in registers; 1 = off, 0 = 8 */
    uint32_t waves;        /* waves per workgroup of the count kernel: 2, 4 or 8 (workgroup
                              size 128, 256 or 512), 0 = 4                              */
    uint32_t reserved[1];
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

/* Same count (over owned rows [0, n_owned)) of a graph that also has
 * target-side in-lists, from lstc_clean_orient (whole graphs) or lstc_in_lists
 * (any CSR, owned sources only): every arc (u, w) streams the shorter of N+(u)
 * and N+(w) through an LDS table of the other, and rows with d+ <= tiny_max
 * compare both lists lane-per-arc, which on power-law graphs streams well under
 * half the entries of lstc_count.  d_col_seg / d_in_seg: the graph's segment
 * words from lstc_segments, or both NULL (then each segment's list is located
 * through d_off).  Workspace: lstc_count_workspace(n_rows, ., m). */
lstc_status lstc_count_split(const uint64_t *d_off, const uint32_t *d_col, const uint64_t *d_in_off,
                             const uint32_t *d_in_col, const uint64_t *d_col_seg, const uint64_t *d_in_seg,
                             uint32_t n_rows, uint32_t n_owned, uint64_t m, uint64_t *d_count,
                             const lstc_tune *tune, void *d_ws, size_t ws_bytes, lstc_stream s);

/* The row-aligned streaming copy for a CSR whose in-lists and segment words
 * came from lstc_in_lists_segments (halo parts, a CSR read from a file): every
 * row of d_off / d_col again in d_scol starting on a 32-entry boundary, and the
 * start field of every segment word with a list rewritten in place to point
 * there: d_col_seg[0 .. n_col_words) (the entries of d_col that have words:
 * off[n_owned] of them for a halo part, m for a whole graph) and
 * d_in_seg[0 .. n_in) (ids d_in_col).  d_scol holds at least m + 31 n_rows
 * entries (< 2^32); *scol_out receives the entries used (the stream is
 * synchronised).  The words then go with d_scol to lstc_count_split_stream or
 * a batch descriptor.  No counterpart in the reference.                      */
lstc_status lstc_stream_copy_workspace(uint32_t n_rows, size_t *bytes);
lstc_status lstc_stream_copy(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint64_t m,
                             uint64_t n_col_words, uint64_t *d_col_seg, const uint32_t *d_in_col, uint64_t n_in,
                             uint64_t *d_in_seg, uint32_t *d_scol, uint64_t scol_cap, uint64_t *scol_out,
                             void *d_ws, size_t ws_bytes, lstc_stream s);

/* lstc_count_split over the outputs of lstc_clean_orient_stream: the lists
 * named by the segment words are read from the streaming copy d_scol
 * (scol_len entries), tables and everything else from d_off / d_col.  Same
 * count, same workspace (lstc_count_workspace).  Replaces the same reference
 * code as lstc_count_split.                                                  */
lstc_status lstc_count_split_stream(const uint64_t *d_off, const uint32_t *d_col, const uint64_t *d_in_off,
                                    const uint32_t *d_in_col, const uint64_t *d_col_seg, const uint64_t *d_in_seg,
                                    const uint32_t *d_scol, uint64_t scol_len, uint32_t n_rows, uint32_t n_owned,
                                    uint64_t m, uint64_t *d_count, const lstc_tune *tune, void *d_ws,
                                    size_t ws_bytes, lstc_stream s);

/* lstc_count_split without a workspace argument, for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  The
 * engine's scratch is a static device array of lstc_count_split_static_limit()
 * bytes (32 MiB) per GPU, held until the process ends; a graph whose scratch
 * need (lstc_count_workspace) exceeds it is refused with LSTC_ERR_INVALID.
 * The call allocates nothing.  Calls that overlap on one GPU share the array
 * and race: keep them on one stream or serialise them.                      */
size_t lstc_count_split_static_limit(void);
lstc_status lstc_count_split_static(const uint64_t *d_off, const uint32_t *d_col, const uint64_t *d_in_off,
                                    const uint32_t *d_in_col, const uint64_t *d_col_seg, const uint64_t *d_in_seg,
                                    uint32_t n_rows, uint32_t n_owned, uint64_t m, uint64_t *d_count,
                                    const lstc_tune *tune, lstc_stream s);

/* Per-edge triangle support and per-vertex triangle counts (local clustering,
 * k-truss), the step after the reference's per-source partials g_sum[u]
 * (src/WithoutPartition/ShareKernelTCV62_update_v1.cu:175-186).  Over the
 * triangles whose source row is owned (u < n_owned; all of them when
 * n_owned = n_rows): d_support[m] (uint32, may be NULL) receives for arc j of
 * the oriented CSR the number of triangles containing that edge, d_per_vertex
 * [n_rows] (uint64, may be NULL) the number containing each vertex.  Each sums
 * to 3x the count.  m < 2^32 - 4096 (32-bit positions inside the walk, as
 * lstc_count).  Stream-ordered, no workspace.                               */
lstc_status lstc_triangle_support(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint32_t n_owned,
                                  uint64_t m, uint32_t *d_support, uint64_t *d_per_vertex, lstc_stream s);

/* The same outputs from the split rule of lstc_count_split: with the graph's
 * target-side in-lists (lstc_in_lists, or the preprocessing's), every arc
 * streams its shorter list.  Identical results; the workspace holds the work
 * items (lstc_triangle_support_split_workspace).                             */
lstc_status lstc_triangle_support_split_workspace(uint32_t n_rows, uint64_t m, size_t *bytes);
lstc_status lstc_triangle_support_split(const uint64_t *d_off, const uint32_t *d_col, const uint64_t *d_in_off,
                                        const uint32_t *d_in_col, uint32_t n_rows, uint32_t n_owned, uint64_t m,
                                        uint32_t *d_support, uint64_t *d_per_vertex, void *d_ws, size_t ws_bytes,
                                        lstc_stream s);

/* Target-side in-lists of an oriented CSR for split counting: for every row w,
 * the owned sources u < n_owned of arcs u -> w with d+(u) <= d+(w) (row lengths
 * of this CSR), ascending.  d_in_off[n_rows+1], d_in_col[m] (at most m entries);
 * *n_entries receives the entry count (the stream is synchronised).  With
 * n_owned = n_rows this equals the in-lists lstc_clean_orient emits.  Enables
 * split counting of the halo subgraphs of a partition (owned rows first, as
 * src/WithPartition/without_sorting_tc.cu:400 counts them).                  */
lstc_status lstc_in_lists_workspace(uint32_t n_rows, uint32_t n_owned, uint64_t m, size_t *bytes);
lstc_status lstc_in_lists(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint32_t n_owned,
                          uint64_t m, uint64_t *d_in_off, uint32_t *d_in_col, uint64_t *n_entries, void *d_ws,
                          size_t ws_bytes, lstc_stream s);

/* lstc_in_lists and the segment words the split count reads, in one call and
 * the same passes over the arcs (the in-list pass writes the column words, the
 * in_col pass the in-list words), so a bare oriented CSR -- the reference's
 * input, src/WithoutPartition/ShareKernelTCV62_update_v1.cu:207-230 -- is made
 * ready for lstc_count_split with one gather of row offsets per entry.  Words
 * are written for every in-list entry (d_in_seg) and for the d_col entries of
 * owned rows (d_col_seg; rows < n_owned come first, so these are positions
 * [0, d_off[n_owned]): every entry for a whole graph; the count reads no other
 * column word).  In-list entries are arcs of owned rows, so d_in_col, d_col_seg
 * and d_in_seg need room for d_off[n_owned] entries (m for a whole graph).
 * m < 2^32 - 4096.  Workspace: lstc_in_lists_workspace.  The stream is
 * synchronised.                                                              */
lstc_status lstc_in_lists_segments(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint32_t n_owned,
                                   uint64_t m, uint64_t *d_in_off, uint32_t *d_in_col, uint64_t *d_col_seg,
                                   uint64_t *d_in_seg, uint64_t *n_entries, void *d_ws, size_t ws_bytes,
                                   lstc_stream s);

/* Segment words of a graph prepared for split counting: for every entry v of
 * d_col[m] and of d_in_col[n_in] (n_in = d_in_off[n_rows]), the word
 * start | length << 32 of v's row (start = d_off[v], a column index below
 * 2^32; length 0 for ids v >= n_rows).  Built once per graph (one pass over
 * the arcs, stream-ordered, no workspace); the count then reads one coalesced
 * word per segment instead of two row offsets at a random place (the
 * reference's kernel reads row_ptr[v], row_ptr[v+1] per neighbour,
 * src/WithoutPartition/ShareKernelTCV62_update_v1.cu:80-81).  8 bytes per
 * entry: d_col_seg[m], d_in_seg[n_in].                                        */
lstc_status lstc_segments(const uint64_t *d_off, const uint32_t *d_col, uint32_t n_rows, uint64_t m,
                          const uint32_t *d_in_col, uint64_t n_in, uint64_t *d_col_seg, uint64_t *d_in_seg,
                          lstc_stream s);

/* ------------------------------------------------------------------ batching
 * Many graphs (e.g. the k halo parts of a partition, or unrelated graphs of
 * uneven size) counted by ONE persistent launch over one degree-binned ticket
 * queue; d_counts[G] receives each graph's owned-row count.  Replaces the
 * reference's per-part loop of launch + synchronise + reduce + copy
 * (src/WithPartition/without_sorting_tc.cu:244-465).                          */
typedef struct lstc_graph_desc {
    const uint64_t *off;     /* device, n_rows + 1 */
    const uint32_t *col;     /* device, m          */
    uint64_t *per_row;       /* device, n_owned, or NULL */
    const uint64_t *in_off;  /* target-side in-lists (split count) or NULL; used   */
    const uint32_t *in_col;  /* only when per_row is NULL                          */
    uint64_t m;
    uint32_t n_rows;
    uint32_t n_owned;
    const uint64_t *col_seg; /* segment words (lstc_segments) or NULL; used with   */
    const uint64_t *in_seg;  /* the in-lists they describe                        */
    const uint32_t *scol;    /* streaming copy the segment words point into, with  */
    uint64_t scol_len;       /* its entries (lstc_clean_orient_stream), or NULL, 0 */
} lstc_graph_desc;          /* at most 2047 graphs per launch */

lstc_status lstc_count_batch_workspace(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune,
                                       size_t *bytes);
/* Reads the host array `graphs` asynchronously, so it synchronises the stream. */
lstc_status lstc_count_batch(const lstc_graph_desc *graphs, uint32_t G, uint64_t *d_counts,
                             const lstc_tune *tune, void *d_ws, size_t ws_bytes, lstc_stream s);

/* A plan owns device copies of the descriptors and its workspace (allocated at
 * creation), so running it is pure kernel enqueues; launch_graph captures that
 * sequence once into a hipGraph (per d_counts) and replays it.  One plan has
 * one workspace: runs of the same plan must not overlap (use one stream, or
 * one plan per stream).  The graphs' CSRs must stay allocated while the plan
 * is used. */
typedef struct lstc_batch_plan lstc_batch_plan;
lstc_status lstc_batch_plan_create(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune,
                                   lstc_batch_plan **out);
lstc_status lstc_batch_plan_run(lstc_batch_plan *plan, uint64_t *d_counts, lstc_stream s);
lstc_status lstc_batch_plan_launch_graph(lstc_batch_plan *plan, uint64_t *d_counts, lstc_stream s);
lstc_status lstc_batch_plan_destroy(lstc_batch_plan *plan);

/* ------------------------------------------------------------------ work and timing
 * Work of one count of each graph, counted on the device from the graph and
 * the tuning (no reference counterpart; diagnostic for rooflines): the items
 * the planner emits, the entries read to build their LDS tables, the lists
 * located (through segment words, or an id and two row offsets), the list
 * entries streamed, and from those the algorithmic bytes of the count:
 *   4 (hashed_keys + streamed) + 8 or 20 per segment + 20 per tiny_segment
 *   + 16 (items + chunk_items) + 32 tiny_rows + 16 (items + tiny_rows)
 *   + 32 or 16 per planned row (two planner passes over the row offsets).
 * out[G]; workspace from lstc_work_stats_workspace; synchronises the stream. */
typedef struct lstc_work {
    uint64_t rows;          /* rows the planner visits                              */
    uint64_t items;         /* hash-path items: source sides + in-list chunks        */
    uint64_t chunk_items;   /* of which in-list chunks                               */
    uint64_t tiny_rows;     /* rows counted one lane per arc                         */
    uint64_t hashed_keys;   /* entries read to build tables (+ tiny rows' lists)     */
    uint64_t segments;      /* lists located on the hash path                        */
    uint64_t tiny_segments; /* arcs of the lane-per-arc path                         */
    uint64_t streamed;      /* list entries streamed through tables or compared      */
    uint64_t bytes;         /* algorithmic bytes of one count                        */
} lstc_work;
lstc_status lstc_work_stats_workspace(uint32_t G, size_t *bytes);
lstc_status lstc_work_stats(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune,
                            lstc_work *out, void *d_ws, size_t ws_bytes, lstc_stream s);

/* Count-kernel timing (diagnostic; the reference brackets its kernel with
 * cudaEvents, src/WithoutPartition/ShareKernelTCV62_update_v1.cu:263-290).
 * While enabled, every launch of the count kernel on a stream that is not
 * being captured is bracketed by two hipEvents on that stream.
 * lstc_timing_read waits for the recorded launches, returns their summed
 * kernel milliseconds and their number, and forgets them.                    */
lstc_status lstc_timing_enable(int on);
lstc_status lstc_timing_read(double *kernel_ms, uint64_t *launches);

/* ------------------------------------------------------------------ autotuning
 * Replaces the reference's compile-time N_THREADS_PER_BLOCK / SHARED_MEM sweep
 * (src/WithoutPartition/ShareKernelTCV62_update_v1.cu:8-9, src/profiling/occup_*.txt).
 * lstc_autotune times candidate lstc_tune values on the caller's batch
 * (coordinate descent, hipEvents, `reps` launches each) and returns the fastest;
 * every candidate must reproduce the default counts exactly.  It allocates its
 * own scratch and synchronises the stream.  The cache is a TSV of
 * (device, shape class, knobs, ms); the last matching line wins.              */
lstc_status lstc_shape_class(const lstc_graph_desc *graphs, uint32_t G, uint32_t *shape_class);
/* The tuning a launch over these graphs runs with: defaults filled in, ranges
 * clamped, and the small-launch default of small_max applied (host only, reads
 * only the descriptors' sizes).  A tune from lstc_autotune resolves to itself. */
lstc_status lstc_tune_resolve(const lstc_graph_desc *graphs, uint32_t G, const lstc_tune *tune, lstc_tune *out);
lstc_status lstc_autotune(const lstc_graph_desc *graphs, uint32_t G, uint32_t reps, lstc_tune *best,
                          double *best_ms, lstc_stream s);
lstc_status lstc_tune_cache_lookup(const char *path, uint32_t shape_class, lstc_tune *out);
lstc_status lstc_tune_cache_store(const char *path, uint32_t shape_class, const lstc_tune *tune, double ms);

/* ------------------------------------------------------------------ serving
 * No reference counterpart (the reference runs one process per graph).  A server
 * owns `hbm_budget_bytes` of HBM (allocated once at creation): a pool for the
 * requests' working sets and resident CSRs plus a reserve for batch work queues.
 * Requests are admitted FIFO as memory frees up; admitted requests are
 * generated and cleaned/oriented on a prep stream while the previous batch is
 * counted on a count stream; every prepared request joins the next persistent
 * batched launch (continuous batching).  A request whose working set exceeds
 * the budget is refused at submission with LSTC_ERR_OVER_BUDGET.
 * Latency = submission to result available on the host.                     */
#define LSTC_REQ_RMAT 0u /* generate RMAT(scale, m_raw, a, b, c, seed), clean, orient */
#define LSTC_REQ_ER 1u   /* generate n vertices, m_raw uniform pairs, seed, clean, orient */
#define LSTC_REQ_CSR 2u  /* count a caller-resident oriented CSR (off, col, n_rows, m) */

typedef struct lstc_request {
    uint32_t kind;
    uint32_t scale;      /* RMAT: n = 2^scale                        */
    uint32_t n;          /* ER: vertices                             */
    uint32_t n_rows;     /* CSR: rows (all owned)                    */
    uint64_t m_raw;      /* RMAT / ER: raw edge samples              */
    uint64_t seed;
    double a, b, c;      /* RMAT quadrant probabilities              */
    const uint64_t *off; /* CSR: device offsets, valid until the result is taken */
    const uint32_t *col; /* CSR: device ids                          */
    uint64_t m;          /* CSR: arcs                                */
} lstc_request;

typedef struct lstc_result {
    uint64_t count;    /* exact triangle count                        */
    uint64_t edges;    /* undirected edges of the counted graph        */
    double latency_s;  /* submission -> result on the host             */
} lstc_result;

typedef struct lstc_server_stats {
    uint64_t submitted, completed, rejected, batches, max_batch_graphs;
    uint64_t pool_bytes, pool_used_bytes, edges, triangles;
    double p50_s, p99_s, mean_s, max_s;
    uint64_t prep_groups, max_prep_graphs; /* preparations (groups of requests) and the largest */
} lstc_server_stats;

typedef struct lstc_server lstc_server;
lstc_status lstc_server_create(size_t hbm_budget_bytes, uint32_t max_batch, const lstc_tune *tune,
                               lstc_server **out);
lstc_status lstc_server_submit(lstc_server *srv, const lstc_request *req, uint64_t *id);
/* LSTC_ERR_NOT_READY until the result exists; a taken result is removed. */
lstc_status lstc_server_poll(lstc_server *srv, uint64_t id, lstc_result *out);
lstc_status lstc_server_wait(lstc_server *srv, uint64_t id, double timeout_s, lstc_result *out);
/* Cancels a request that has not been admitted yet: it is dropped before it
 * takes memory and its result is LSTC_ERR_CANCELLED.  LSTC_ERR_NOT_READY for a
 * request already admitted or finished (it keeps its normal result),
 * LSTC_ERR_INVALID for an id this server never returned. */
lstc_status lstc_server_cancel(lstc_server *srv, uint64_t id);
lstc_status lstc_server_get_stats(lstc_server *srv, lstc_server_stats *out);
/* Finishes every submitted request, then frees the server. */
lstc_status lstc_server_destroy(lstc_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* LSTC_H */
