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
    const double t0 = now_s();
    uint64_t *d_off = upload(g.off);
    uint32_t *d_col = upload(g.col);
    uint32_t *d_part = nullptr;
    CLI_HIP(hipMalloc((void **)&d_part, sizeof(uint32_t) * (g.n ? g.n : 1)));
    if (lstc_partition_read(argv[3], (uint32_t)g.n, (uint32_t)k, d_part, nullptr) != LSTC_OK) {
        fprintf(stderr, "Error opening file: %s (%s)\n", argv[3], lstc_last_error());
        return 1;
    }
    Workspace ws;
    size_t need = 0;
    std::vector<uint32_t> nl(k), no(k);
    std::vector<uint64_t> ml(k);
    std::vector<lstc_halo_out> outs(k);
    auto alloc_out = [&](uint64_t p) {
        lstc_halo_out &o = outs[p];
        o = lstc_halo_out{};
        CLI_HIP(hipMalloc((void **)&o.off, sizeof(uint64_t) * ((size_t)nl[p] + 1)));
        CLI_HIP(hipMalloc((void **)&o.col, sizeof(uint32_t) * (ml[p] ? ml[p] : 1)));
        o.n_local_cap = nl[p];
        o.m_local_cap = ml[p];
    };
    if (k <= 64) {  // every part at once: each pass over the graph serves all k parts
        CLI_CHECK(lstc_halo_parts_workspace((uint32_t)g.n, (uint32_t)k, &need));
        void *w = ws.get(need);
        CLI_CHECK(lstc_halo_parts_sizes(d_off, d_col, (uint32_t)g.n, d_part, (uint32_t)k, mode, nl.data(), no.data(),
                                        ml.data(), w, ws.bytes, nullptr));
        for (uint64_t p = 0; p < k; ++p) alloc_out(p);
        CLI_CHECK(lstc_halo_parts_extract(d_off, d_col, (uint32_t)g.n, d_part, (uint32_t)k, mode, outs.data(), w,
                                          ws.bytes, nullptr));
    } else {  // one part per call
        CLI_CHECK(lstc_halo_workspace((uint32_t)g.n, &need));
        void *w = ws.get(need);
        for (uint64_t p = 0; p < k; ++p) {
            CLI_CHECK(lstc_halo_sizes(d_off, d_col, (uint32_t)g.n, d_part, (uint32_t)p, mode, &nl[p], &no[p], &ml[p],
                                      w, ws.bytes, nullptr));
            alloc_out(p);
            CLI_CHECK(lstc_halo_extract(d_off, d_col, (uint32_t)g.n, d_part, (uint32_t)p, mode, outs[p].off,
                                        outs[p].col, nullptr, nl[p], ml[p], w, ws.bytes, nullptr));
        }
    }
    std::vector<lstc_text::PartCsr> parts(k);
    for (uint64_t p = 0; p < k; ++p) {
        parts[p].n_local = nl[p];
        parts[p].n_owned = no[p];
        parts[p].off = download(outs[p].off, (size_t)nl[p] + 1);
        parts[p].col = download(outs[p].col, ml[p]);
        (void)hipFree(outs[p].off);
        (void)hipFree(outs[p].col);
        printf("%llu th Partitions: owned %u, with halo %u, arcs %llu\n", (unsigned long long)p, no[p], nl[p],
               (unsigned long long)ml[p]);
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
