/* Host driver for the sanitizer build of the oracle (oracle/lstc_oracle.c): runs the
 * whole CPU path on small seeded graphs and checks that its independent pieces
 * agree -- clean -> orient -> merge counter == block-per-row restatement ==
 * support sums / 3 == sum over halo parts, for both halo modes.  Exit code 0
 * when all agree; any out-of-bounds access or undefined behaviour is reported
 * by the sanitizers and ends the run. */
#include <stdio.h>
#include <stdlib.h>

#include "lstc_oracle.h"

static int check_graph(const char *name, uint32_t n, uint64_t m_raw, const uint32_t *s, const uint32_t *d) {
    uint64_t *soff = malloc(8 * ((size_t)n + 1));
    uint32_t *scol = malloc(4 * (2 * m_raw + 1));
    uint32_t n_out = 0;
    uint64_t m = 0;
    lo_clean(n, m_raw, s, d, 1, &n_out, &m, soff, scol);
    uint64_t *off = malloc(8 * ((size_t)n_out + 1));
    uint32_t *col = malloc(4 * (m + 1));
    lo_orient(n_out, soff, scol, off, col);
    uint64_t *per_row = calloc((size_t)n_out + 1, 8), *g_sum = calloc((size_t)n_out + 1, 8);
    const uint64_t t = lo_count_merge(n_out, n_out, off, col, per_row, 2);
    const uint64_t t1 = lo_count_rows_k1(n_out, n_out, off, col, g_sum);
    uint32_t *sup = calloc(m + 1, 4);
    uint64_t *pv = calloc((size_t)n_out + 1, 8);
    lo_triangle_support(n_out, n_out, off, col, sup, pv);
    uint64_t ss = 0, sv = 0;
    for (uint64_t i = 0; i < m; ++i) ss += sup[i];
    for (uint32_t v = 0; v < n_out; ++v) sv += pv[v];
    int bad = t != t1 || ss != 3 * t || sv != 3 * t;
    for (uint32_t v = 0; v < n_out; ++v) bad |= per_row[v] != g_sum[v];
    lo_stats_t st;
    lo_stats(n_out, off, col, &st);
    bad |= st.m != m;
    uint32_t *part = malloc(4 * ((size_t)n_out + 1));
    const uint32_t k = 5;
    lo_partition_random(n_out, k, 7, part);
    for (int mode = 0; mode < 2; ++mode) {
        uint64_t sum = 0;
        for (uint32_t p = 0; p < k; ++p) {
            uint32_t nl = 0, no = 0;
            uint64_t ml = 0;
            lo_halo_extract(n_out, soff, scol, part, p, mode, &nl, &no, &ml, NULL, NULL, NULL);
            uint64_t *ho = malloc(8 * ((size_t)nl + 1));
            uint32_t *hc = malloc(4 * (ml + 1)), *l2g = malloc(4 * ((size_t)nl + 1));
            lo_halo_extract(n_out, soff, scol, part, p, mode, &nl, &no, &ml, ho, hc, l2g);
            sum += lo_count_merge(nl, no, ho, hc, NULL, 1);
            free(ho); free(hc); free(l2g);
        }
        bad |= sum != t;
    }
    printf("%s: n=%u m=%llu triangles=%llu %s\n", name, n_out, (unsigned long long)m, (unsigned long long)t,
           bad ? "MISMATCH" : "ok");
    free(soff); free(scol); free(off); free(col); free(per_row); free(g_sum); free(sup); free(pv); free(part);
    return bad;
}

int main(void) {
    int bad = 0;
    const uint64_t m = 20000;
    uint32_t *s = malloc(4 * m * 3), *d = malloc(4 * m * 3);
    lo_gen_rmat(11, m, 0.57, 0.19, 0.19, 3, s, d);
    bad |= check_graph("rmat s11", 1u << 11, m, s, d);
    lo_gen_er(3000, m, 4, s, d);
    bad |= check_graph("er", 3000, m, s, d);
    if (lo_gen_chung_lu(5000, 2.1, 8.0, 500.0, m, 5, s, d) == 0) bad |= check_graph("chung-lu", 5000, m, s, d);
    lo_gen_grid(60, 80, 0.3, 6, s, d);
    bad |= check_graph("grid", 60 * 80, 3ull * 60 * 80, s, d);
    bad |= lo_grid_num_diagonals(60, 80, 0.3, 6) * 2 == 0;
    bad |= check_graph("empty", 10, 0, s, d);
    free(s); free(d);
    return bad;
}
