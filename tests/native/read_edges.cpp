// Host driver for the CPU tests of the edge-list readers (trianglecount_amd/cli/lstc_text.h):
// prints "n m" then one "src dst" line per raw edge, or "error: <message>" and exits 1.
#include "lstc_text.h"

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    uint64_t n = 0;
    std::vector<uint32_t> s, d;
    std::string err;
    if (!lstc_text::read_edge_list(argv[1], n, s, d, err)) {
        printf("error: %s\n", err.c_str());
        return 1;
    }
    printf("%llu %zu\n", (unsigned long long)n, s.size());
    for (size_t i = 0; i < s.size(); ++i) printf("%u %u\n", s[i], d[i]);
    return 0;
}
