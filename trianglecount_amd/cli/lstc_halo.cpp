// lstc_halo <k> <part.csr> <node_parts> [--prune] -- halo subgraphs of a k-way partition.
//
// Drop-in for the reference's round_clean_sg (README.md:39-44;
// preprocess_GPU/Halo_cleanned_preprocess/round_clean_subgraph.cu:904-976):
// reads the symmetric CSR with its degree line and the node_parts vector, writes
// "<node_parts>_output.csr" in the reference's partitioned format (owned rows
// first, then halo rows; rows ascending).  Default: the un-pruned induced
// subgraph on P u N(P); --prune keeps P u N+(P), which preserves every count.
#include <string.h>

#include <string>

#include "lstc_cli.h"
#include "lstc_text.h"

using namespace lstc_cli;

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "Usage: %s <nopart> <filename> <node_parts> [--prune]\n", argv[0]);
        return 1;
    }
    const uint64_t k = strtoull(argv[1], nullptr, 10);
    const int mode = (argc > 4 && strcmp(argv[4], "--prune") == 0) ? 1 : 0;
    lstc_text::Csr g;
    std::vector<uint64_t> deg;
    std::string err;
    if (!lstc_text::read_part_input(argv[2], g, deg, err)) {
        fprintf(stderr, "Error opening file: %s (%s)\n", argv[2], err.c_str());
        return 1;
    }
    // the orientation uses the global degree; it must be the symmetric row length
    for (uint64_t v = 0; v < g.n && v < deg.size(); ++v)
        if (deg[v] != g.off[v + 1] - g.off[v]) {
            fprintf(stderr, "in_deg[%llu] differs from its row length: unsupported input\n", (unsigned long long)v);
            return 1;
        }
    std::vector<uint32_t> part;
    if (!lstc_text::read_node_parts(argv[3], g.n, part, err)) {
        fprintf(stderr, "Error opening file: %s (%s)\n", argv[3], err.c_str());
        return 1;
    }
    const double t0 = now_s();
    uint64_t *d_off = upload(g.off);
    uint32_t *d_col = upload(g.col);
    uint32_t *d_part = upload(part);
    Workspace ws;
    size_t need = 0;
    CLI_CHECK(lstc_halo_workspace((uint32_t)g.n, &need));
    void *w = ws.get(need);
    std::vector<lstc_text::PartCsr> parts(k);
    for (uint64_t p = 0; p < k; ++p) {
        uint32_t nl = 0, no = 0;
        uint64_t ml = 0;
        CLI_CHECK(lstc_halo_sizes(d_off, d_col, (uint32_t)g.n, d_part, (uint32_t)p, mode, &nl, &no, &ml, w, ws.bytes,
                                  nullptr));
        uint64_t *o = nullptr;
        uint32_t *c = nullptr;
        CLI_HIP(hipMalloc((void **)&o, sizeof(uint64_t) * ((size_t)nl + 1)));
        CLI_HIP(hipMalloc((void **)&c, sizeof(uint32_t) * (ml ? ml : 1)));
        CLI_CHECK(lstc_halo_extract(d_off, d_col, (uint32_t)g.n, d_part, (uint32_t)p, mode, o, c, nullptr, nl, ml, w,
                                    ws.bytes, nullptr));
        parts[p].n_local = nl;
        parts[p].n_owned = no;
        parts[p].off = download(o, (size_t)nl + 1);
        parts[p].col = download(c, ml);
        (void)hipFree(o);
        (void)hipFree(c);
        printf("%llu th Partitions: owned %u, with halo %u, arcs %llu\n", (unsigned long long)p, no, nl,
               (unsigned long long)ml);
    }
    const double t1 = now_s();
    const std::string out = std::string(argv[3]) + "_output.csr";
    if (!lstc_text::write_partitioned(out.c_str(), g.n, g.header_edges, parts)) {
        fprintf(stderr, "Error opening file: %s\n", out.c_str());
        return 1;
    }
    printf("Total EXE Time: %f seconds.\n", t1 - t0);
    printf("Total Time measured:%f seconds.\n", now_s() - t0);
    (void)hipFree(d_off);
    (void)hipFree(d_col);
    (void)hipFree(d_part);
    return 0;
}
