// lstc_preprocess <edges.{tsv,txt,mtx}> [--parts k1,k2,..] [--seed S] [--keep-isolated]
//
// GPU replacement for the reference's Python preprocessing:
//   Med_dataset/without_part_preprocess.py (clean + orient, writes <name>_output.csr)
//   preprocess_GPU/generate_clean_undir_dataset/Halo_creation_preprocess.py
//     (writes <name>_part.csr and <name>_part.csr_node_parts_{random,contig}_<k>)
// Loading follows :126-177 (GraphChallenge .tsv, SNAP .txt, SuiteSparse .mtx);
// cleaning and orientation run on the device (lstc_clean_orient).  The random
// assignment is seeded (the reference's np.random.randint is not); the contig
// assignment gives part i the block [i*l, (i+1)*l), l = n // k, the remainder
// staying in part 0 as in :292-300.
#include <string.h>

#include <string>

#include "lstc_cli.h"
#include "lstc_text.h"

using namespace lstc_cli;

static std::string stem(const char *path) {
    std::string p(path);
    const size_t slash = p.find_last_of('/');
    if (slash != std::string::npos) p = p.substr(slash + 1);
    const size_t dot = p.find_last_of('.');
    if (dot != std::string::npos) p = p.substr(0, dot);
    return p;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <edges.tsv|.txt|.mtx> [--parts 2,4,8] [--seed S] [--keep-isolated]\n", argv[0]);
        return 1;
    }
    std::vector<uint32_t> ks;
    uint64_t seed = 1;
    int compact = 1;
    for (int i = 2; i < argc; ++i) {
        if (strcmp(argv[i], "--parts") == 0 && i + 1 < argc) {
            std::string s(argv[++i]);
            size_t a = 0;
            while (a < s.size()) {
                size_t b = s.find(',', a);
                if (b == std::string::npos) b = s.size();
                ks.push_back((uint32_t)strtoul(s.substr(a, b - a).c_str(), nullptr, 10));
                a = b + 1;
            }
        } else if (strcmp(argv[i], "--seed") == 0 && i + 1 < argc) {
            seed = strtoull(argv[++i], nullptr, 10);
        } else if (strcmp(argv[i], "--keep-isolated") == 0) {
            compact = 0;
        }
    }
    uint64_t n = 0;
    std::vector<uint32_t> src, dst;
    std::string err;
    const double t0 = now_s();
    if (!lstc_text::read_edge_list(argv[1], n, src, dst, err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    printf("Data Loading Successfull!!!! \tTime Taken of Loading is : %.4f Seconds\n", now_s() - t0);
    const uint64_t m_raw = src.size();
    uint32_t *d_src = upload(src), *d_dst = upload(dst);
    uint64_t *d_off = nullptr, *d_soff = nullptr;
    uint32_t *d_col = nullptr, *d_scol = nullptr;
    CLI_HIP(hipMalloc((void **)&d_off, sizeof(uint64_t) * (n + 1)));
    CLI_HIP(hipMalloc((void **)&d_soff, sizeof(uint64_t) * (n + 1)));
    CLI_HIP(hipMalloc((void **)&d_col, sizeof(uint32_t) * (m_raw ? m_raw : 1)));
    CLI_HIP(hipMalloc((void **)&d_scol, sizeof(uint32_t) * (2 * m_raw ? 2 * m_raw : 1)));
    Workspace ws;
    size_t need = 0;
    CLI_CHECK(lstc_clean_orient_workspace((uint32_t)n, m_raw, &need));
    void *w = ws.get(need);
    uint32_t n_out = 0;
    uint64_t m_out = 0;
    const double t1 = now_s();
    CLI_CHECK(lstc_clean_orient(d_src, d_dst, m_raw, (uint32_t)n, compact, d_soff, d_scol, d_off, d_col, nullptr,
                                nullptr, &n_out, &m_out, w, ws.bytes, nullptr));
    printf("Graph Construction + Converting Successfull!!!! \tTime Taken : %.4f Seconds (vertices %u, edges %llu)\n",
           now_s() - t1, n_out, (unsigned long long)m_out);
    const std::vector<uint64_t> off = download(d_off, (size_t)n_out + 1), soff = download(d_soff, (size_t)n_out + 1);
    const std::vector<uint32_t> col = download(d_col, m_out), scol = download(d_scol, 2 * m_out);
    const std::string name = stem(argv[1]);
    const std::string o1 = name + "_output.csr";
    // header edges = directed edges of the symmetrised graph, as DGL's num_edges() reports
    lstc_text::write_oriented(o1.c_str(), n_out, 2 * m_out, off.data(), n_out, col.data(), m_out);
    printf("wrote %s\n", o1.c_str());
    if (!ks.empty()) {
        const std::string o2 = name + "_part.csr";
        lstc_text::write_part_input(o2.c_str(), soff.data(), n_out, scol.data(), 2 * m_out);
        printf("wrote %s\n", o2.c_str());
        uint32_t *d_part = nullptr;
        CLI_HIP(hipMalloc((void **)&d_part, sizeof(uint32_t) * (n_out ? n_out : 1)));
        for (uint32_t k : ks) {
            if (k == 0) continue;
            CLI_CHECK(lstc_partition_random(n_out, k, seed, d_part, nullptr));
            const std::vector<uint32_t> part = download(d_part, n_out);
            const std::string pr = o2 + "_node_parts_random_" + std::to_string(k);
            lstc_text::write_node_parts(pr.c_str(), part.data(), n_out);
            CLI_CHECK(lstc_partition_contig(n_out, k, d_part, nullptr));
            const std::vector<uint32_t> contig = download(d_part, n_out);
            const std::string pc = o2 + "_node_parts_contig_" + std::to_string(k);
            lstc_text::write_node_parts(pc.c_str(), contig.data(), n_out);
            printf("wrote %s, %s\n", pr.c_str(), pc.c_str());
        }
        (void)hipFree(d_part);
    }
    printf("Data Preprocessing is Successfull Total Time taken: %.4f Seconds\n", now_s() - t0);
    (void)hipFree(d_src); (void)hipFree(d_dst); (void)hipFree(d_off); (void)hipFree(d_soff);
    (void)hipFree(d_col); (void)hipFree(d_scol);
    return 0;
}
