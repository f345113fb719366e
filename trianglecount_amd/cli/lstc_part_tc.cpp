// lstc_part_tc <file> <k> [--per-part] -- exact count of a partitioned CSR file.
//
// Drop-in for the reference's wo_sort_tc_256 (README.md:57-64;
// src/WithPartition/without_sorting_tc.cu:220-470): same input format, same
// per-part and total lines.  By default all k parts are counted by ONE
// persistent batched launch; --per-part reproduces the reference's loop of one
// launch + synchronise + copy per part.  Each part gets owned-source in-lists
// before the timed count (split counting); --plain counts the CSRs without them.
#include <string.h>

#include "lstc_cli.h"
#include "lstc_text.h"

using namespace lstc_cli;

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <partitioned.csr> <k> [--per-part] [--plain]\n", argv[0]);
        return 1;
    }
    const uint64_t k = strtoull(argv[2], nullptr, 10);
    const bool per_part = has_flag(argc, argv, 3, "--per-part");
    const bool plain = has_flag(argc, argv, 3, "--plain");
    uint64_t V = 0, E = 0;
    std::vector<lstc_text::PartCsr> parts;
    std::string err;
    if (!lstc_text::read_partitioned(argv[1], k, V, E, parts, err)) {
        printf("\nFile Not Opened.........");
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    const double t0 = now_s();
    std::vector<lstc_graph_desc> descs(k);
    std::vector<uint64_t *> offs(k);
    std::vector<uint32_t *> cols(k);
    std::vector<SplitLists> split(k);
    Workspace ws_in;
    for (uint64_t p = 0; p < k; ++p) {
        offs[p] = upload(parts[p].off);
        cols[p] = upload(parts[p].col);
        if (!plain)
            split[p] = build_in_lists(offs[p], cols[p], (uint32_t)parts[p].n_local, (uint32_t)parts[p].n_owned,
                                      parts[p].col.size(), ws_in);
        descs[p] = lstc_graph_desc{offs[p], cols[p], nullptr, split[p].in_off, split[p].in_col, parts[p].col.size(),
                                   (uint32_t)parts[p].n_local, (uint32_t)parts[p].n_owned, split[p].col_seg,
                                   split[p].in_seg};
    }
    uint64_t *d_counts = nullptr;
    CLI_HIP(hipMalloc((void **)&d_counts, sizeof(uint64_t) * (k ? k : 1)));
    Workspace ws;
    hipEvent_t e0, e1;
    CLI_HIP(hipEventCreate(&e0));
    CLI_HIP(hipEventCreate(&e1));
    std::vector<uint64_t> counts(k, 0);
    double kernel_s = 0;
    if (per_part) {
        for (uint64_t p = 0; p < k; ++p) {
            size_t need = 0;
            CLI_CHECK(lstc_count_workspace(descs[p].n_rows, descs[p].n_owned, descs[p].m, nullptr, &need));
            void *w = ws.get(need);
            CLI_HIP(hipEventRecord(e0, nullptr));
            if (plain)
                CLI_CHECK(lstc_count(descs[p].off, descs[p].col, descs[p].n_rows, descs[p].n_owned, descs[p].m,
                                     d_counts + p, nullptr, nullptr, w, ws.bytes, nullptr));
            else
                CLI_CHECK(lstc_count_split(descs[p].off, descs[p].col, descs[p].in_off, descs[p].in_col,
                                           descs[p].col_seg, descs[p].in_seg, descs[p].n_rows, descs[p].n_owned,
                                           descs[p].m, d_counts + p, nullptr, w, ws.bytes, nullptr));
            CLI_HIP(hipEventRecord(e1, nullptr));
            CLI_HIP(hipEventSynchronize(e1));
            float ms = 0;
            CLI_HIP(hipEventElapsedTime(&ms, e0, e1));
            kernel_s += ms / 1e3;
            CLI_HIP(hipMemcpy(&counts[p], d_counts + p, sizeof(uint64_t), hipMemcpyDeviceToHost));
        }
    } else if (k) {
        size_t need = 0;
        CLI_CHECK(lstc_count_batch_workspace(descs.data(), (uint32_t)k, nullptr, &need));
        void *w = ws.get(need);
        CLI_HIP(hipEventRecord(e0, nullptr));
        CLI_CHECK(lstc_count_batch(descs.data(), (uint32_t)k, d_counts, nullptr, w, ws.bytes, nullptr));
        CLI_HIP(hipEventRecord(e1, nullptr));
        CLI_HIP(hipEventSynchronize(e1));
        float ms = 0;
        CLI_HIP(hipEventElapsedTime(&ms, e0, e1));
        kernel_s = ms / 1e3;
        counts = download(d_counts, k);
    }
    uint64_t total = 0;
    for (uint64_t p = 0; p < k; ++p) {
        printf("\nPart %llu executing :", (unsigned long long)p);
        printf("\t%llu\n", (unsigned long long)counts[p]);
        total += counts[p];
    }
    const double t_all = now_s() - t0;
    printf("\nTotal Vertex : %llu ", (unsigned long long)V);
    printf("\tTotal Edges : %llu ", (unsigned long long)(E / 2));
    printf("\tTotal Triangle : %llu ", (unsigned long long)total);
    printf("\tTotal Kernel Time : %.4f sec", kernel_s);
    printf("\tTotal Time : %f \n\n", t_all);
    for (uint64_t p = 0; p < k; ++p) {
        (void)hipFree(offs[p]);
        (void)hipFree(cols[p]);
        split[p].release();
    }
    (void)hipFree(d_counts);
    return 0;
}
