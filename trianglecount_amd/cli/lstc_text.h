// lstc_text.h -- readers/writers for the reference's text graph formats and raw
// edge lists, shared by the lstc_* command-line tools (host C++ only).
//
// Formats (reference files, relative to its root):
//   oriented CSR  "<name>_output.csr": `V E |row_ptr| |col|` then row_ptr then col
//                 (writer Med_dataset/without_part_preprocess.py:298-311,
//                  reader src/WithoutPartition/ShareKernelTCV62_update_v1.cu:207-230)
//   partition input "<name>_part.csr": `V E` / `V V+1 |col| 0` / row_ptr / col / in_deg
//                 (writer preprocess_GPU/generate_clean_undir_dataset/Halo_creation_preprocess.py:259-278,
//                  reader preprocess_GPU/Halo_cleanned_preprocess/round_clean_subgraph.cu:924-952)
//   node parts    one integer per vertex (Halo_creation_preprocess.py:308-312; reader :964-968)
//   partitioned CSR "<node_parts>_output.csr": `V E`, then per part
//                 `Num_Nodes Num_Nodes+1 Num_Edges |P|` / row_ptr / col
//                 (writer round_clean_subgraph.cu:371,850-865; reader src/WithPartition/without_sorting_tc.cu:242-286)
//   edge lists    .tsv (GraphChallenge: src dst weight), .txt (SNAP, 4 header lines),
//                 .mtx / .mmio (Matrix Market coordinate, 1-based)
//                 (loaders Med_dataset/without_part_preprocess.py:126-177)
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

namespace lstc_text {

// Whole-file reader with a fast unsigned integer scanner.
struct Scanner {
    std::vector<char> buf;
    size_t pos = 0;
    bool ok = false;

    explicit Scanner(const char *path) {
        FILE *f = fopen(path, "rb");
        if (!f) return;
        fseek(f, 0, SEEK_END);
        long sz = ftell(f);
        fseek(f, 0, SEEK_SET);
        if (sz < 0) { fclose(f); return; }
        buf.resize((size_t)sz + 1);
        size_t got = fread(buf.data(), 1, (size_t)sz, f);
        fclose(f);
        buf[got] = 0;
        buf.resize(got + 1);
        ok = true;
    }
    // next unsigned integer; skips anything that is not a digit; false at the end
    // of the file or for a number that does not fit 64 bits
    bool next(uint64_t &v) {
        const size_t n = buf.size() - 1;
        while (pos < n && (buf[pos] < '0' || buf[pos] > '9')) ++pos;
        if (pos >= n) return false;
        uint64_t x = 0;
        while (pos < n && buf[pos] >= '0' && buf[pos] <= '9') {
            const uint64_t d = (uint64_t)(buf[pos++] - '0');
            if (x > (UINT64_MAX - d) / 10) return false;
            x = x * 10 + d;
        }
        v = x;
        return true;
    }
    // could the rest of the file hold `count` more integers (a digit and a separator each)?
    bool room(uint64_t count) const { return count <= (buf.size() - pos) / 2 + 1; }
};

// A CSR read from a file is used to index arrays: offsets must start at 0, never
// decrease and end at |col|, and every column id must name a row.
inline bool check_csr(const std::vector<uint64_t> &off, const std::vector<uint32_t> &col, uint64_t rows,
                      std::string &err) {
    if (off.empty() || off[0] != 0) { err = "row_ptr does not start at 0"; return false; }
    for (size_t i = 1; i < off.size(); ++i)
        if (off[i] < off[i - 1]) { err = "row_ptr decreases"; return false; }
    if (off.back() != col.size()) { err = "row_ptr end != |col|"; return false; }
    for (uint32_t c : col)
        if (c >= rows) { err = "column id beyond the rows"; return false; }
    return true;
}

struct Csr {
    uint64_t n = 0, header_edges = 0;  // header_edges: the second header number as written
    std::vector<uint64_t> off;
    std::vector<uint32_t> col;
};

// "<name>_output.csr" (oriented CSR)
inline bool read_oriented(const char *path, Csr &g, std::string &err) {
    Scanner s(path);
    if (!s.ok) { err = std::string("cannot open ") + path; return false; }
    uint64_t V, E, rps, cis;
    if (!s.next(V) || !s.next(E) || !s.next(rps) || !s.next(cis)) { err = "bad header"; return false; }
    if (rps == 0) { err = "empty row_ptr"; return false; }
    if (!s.room(rps) || !s.room(cis) || rps - 1 > 0xFFFFFFFEull) { err = "sizes in the header exceed the file"; return false; }
    g.n = rps - 1;
    g.header_edges = E;
    g.off.resize(rps);
    g.col.resize(cis);
    for (uint64_t i = 0; i < rps; ++i) if (!s.next(g.off[i])) { err = "short row_ptr"; return false; }
    for (uint64_t i = 0; i < cis; ++i) {
        uint64_t v;
        if (!s.next(v)) { err = "short col"; return false; }
        if (v > 0xFFFFFFFFull) { err = "column id beyond 32 bits"; return false; }
        g.col[i] = (uint32_t)v;
    }
    (void)V;
    return check_csr(g.off, g.col, g.n, err);
}

inline bool write_oriented(const char *path, uint64_t V, uint64_t E, const uint64_t *off, uint64_t n,
                           const uint32_t *col, uint64_t m) {
    FILE *f = fopen(path, "w");
    if (!f) return false;
    fprintf(f, "%llu %llu %llu %llu \n", (unsigned long long)V, (unsigned long long)E,
            (unsigned long long)(n + 1), (unsigned long long)m);
    for (uint64_t i = 0; i <= n; ++i) fprintf(f, "%llu ", (unsigned long long)off[i]);
    fprintf(f, "\n");
    for (uint64_t i = 0; i < m; ++i) fprintf(f, "%u ", col[i]);
    fprintf(f, "\n");
    fclose(f);
    return true;
}

// "<name>_part.csr" (symmetric CSR + degrees)
inline bool read_part_input(const char *path, Csr &g, std::vector<uint64_t> &deg, std::string &err) {
    Scanner s(path);
    if (!s.ok) { err = std::string("cannot open ") + path; return false; }
    uint64_t V, E, nm, rps, cis, zero;
    if (!s.next(V) || !s.next(E) || !s.next(nm) || !s.next(rps) || !s.next(cis) || !s.next(zero)) {
        err = "bad header";
        return false;
    }
    if (rps == 0) { err = "empty row_ptr"; return false; }
    if (!s.room(rps) || !s.room(cis) || !s.room(V) || rps - 1 > 0xFFFFFFFEull) {
        err = "sizes in the header exceed the file";
        return false;
    }
    g.n = rps - 1;
    g.header_edges = E;
    g.off.resize(rps);
    g.col.resize(cis);
    for (uint64_t i = 0; i < rps; ++i) if (!s.next(g.off[i])) { err = "short row_ptr"; return false; }
    for (uint64_t i = 0; i < cis; ++i) {
        uint64_t v;
        if (!s.next(v)) { err = "short col"; return false; }
        if (v > 0xFFFFFFFFull) { err = "column id beyond 32 bits"; return false; }
        g.col[i] = (uint32_t)v;
    }
    deg.resize(V);
    for (uint64_t i = 0; i < V; ++i) if (!s.next(deg[i])) { err = "short in_deg"; return false; }
    (void)nm;
    return check_csr(g.off, g.col, g.n, err);
}

inline bool write_part_input(const char *path, const uint64_t *off, uint64_t n, const uint32_t *col, uint64_t m2) {
    FILE *f = fopen(path, "w");
    if (!f) return false;
    fprintf(f, "%llu %llu\n", (unsigned long long)n, (unsigned long long)m2);
    fprintf(f, "%llu %llu %llu 0 \n", (unsigned long long)n, (unsigned long long)(n + 1), (unsigned long long)m2);
    for (uint64_t i = 0; i <= n; ++i) fprintf(f, "%llu ", (unsigned long long)off[i]);
    fprintf(f, "\n");
    for (uint64_t i = 0; i < m2; ++i) fprintf(f, "%u ", col[i]);
    fprintf(f, "\n");
    for (uint64_t i = 0; i < n; ++i) fprintf(f, "%llu ", (unsigned long long)(off[i + 1] - off[i]));
    fprintf(f, "\n");
    fclose(f);
    return true;
}

inline bool read_node_parts(const char *path, uint64_t n, std::vector<uint32_t> &part, std::string &err) {
    Scanner s(path);
    if (!s.ok) { err = std::string("cannot open ") + path; return false; }
    if (!s.room(n)) { err = "short node_parts"; return false; }
    part.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t v;
        if (!s.next(v)) { err = "short node_parts"; return false; }
        part[i] = (uint32_t)v;
    }
    return true;
}

inline bool write_node_parts(const char *path, const uint32_t *part, uint64_t n) {
    FILE *f = fopen(path, "w");
    if (!f) return false;
    for (uint64_t i = 0; i < n; ++i) fprintf(f, "%u ", part[i]);
    fprintf(f, "\n");
    fclose(f);
    return true;
}

struct PartCsr {
    uint64_t n_local = 0, n_owned = 0;
    std::vector<uint64_t> off;
    std::vector<uint32_t> col;
};

// "<node_parts>_output.csr": every part; the reader takes `k` parts as the reference does
inline bool read_partitioned(const char *path, uint64_t k, uint64_t &V, uint64_t &E, std::vector<PartCsr> &parts,
                             std::string &err) {
    Scanner s(path);
    if (!s.ok) { err = std::string("cannot open ") + path; return false; }
    if (!s.next(V) || !s.next(E)) { err = "bad header"; return false; }
    if (!s.room(k)) { err = "fewer parts in the file than asked for"; return false; }
    parts.resize(k);
    for (uint64_t p = 0; p < k; ++p) {
        uint64_t vp, rps, cis, tver;
        if (!s.next(vp) || !s.next(rps) || !s.next(cis) || !s.next(tver)) { err = "bad part header"; return false; }
        if (rps == 0) { err = "empty part row_ptr"; return false; }
        if (!s.room(rps) || !s.room(cis) || rps - 1 > 0xFFFFFFFEull) {
            err = "part sizes exceed the file";
            return false;
        }
        PartCsr &q = parts[p];
        q.n_local = rps - 1;
        q.n_owned = tver;
        if (tver > q.n_local) { err = "more owned rows than rows in a part"; return false; }
        q.off.resize(rps);
        q.col.resize(cis);
        for (uint64_t i = 0; i < rps; ++i) if (!s.next(q.off[i])) { err = "short part row_ptr"; return false; }
        for (uint64_t i = 0; i < cis; ++i) {
            uint64_t v;
            if (!s.next(v)) { err = "short part col"; return false; }
            if (v > 0xFFFFFFFFull) { err = "column id beyond 32 bits"; return false; }
            q.col[i] = (uint32_t)v;
        }
        if (!check_csr(q.off, q.col, q.n_local, err)) return false;
        (void)vp;
    }
    return true;
}

inline bool write_partitioned(const char *path, uint64_t V, uint64_t E, const std::vector<PartCsr> &parts) {
    FILE *f = fopen(path, "w");
    if (!f) return false;
    fprintf(f, "%llu %llu\n", (unsigned long long)V, (unsigned long long)E);
    for (const PartCsr &q : parts) {
        fprintf(f, "%llu %llu %llu %llu\n", (unsigned long long)q.n_local, (unsigned long long)(q.n_local + 1),
                (unsigned long long)q.col.size(), (unsigned long long)q.n_owned);
        for (uint64_t i = 0; i <= q.n_local; ++i) fprintf(f, "%llu ", (unsigned long long)q.off[i]);
        fprintf(f, "\n");
        for (size_t i = 0; i < q.col.size(); ++i) fprintf(f, "%u ", q.col[i]);
        fprintf(f, "\n");
    }
    fclose(f);
    return true;
}

// Raw edge list -> (n, src, dst), as the reference's loaders read it
// (Med_dataset/without_part_preprocess.py:126-177):
//   .tsv          GraphChallenge "src<TAB>dst<TAB>weight", ids as written (:135-144)
//   .txt          SNAP: the first 4 lines are skipped (pandas skiprows=4, :145-154),
//                 then "src<TAB>dst", ids as written
//   .mtx / .mmio  Matrix Market coordinate: '%' lines, one size line
//                 "rows cols nnz", then 1-based "row col [value]" (scipy mmread ->
//                 0-based COO, :126-134, :155-163)
// n = largest id + 1, as dgl.graph((u, v)) counts nodes.  Lines starting with
// '#' or '%' are skipped in every format.  Other extensions are refused, as the
// reference refuses them (:173-175).
inline bool read_edge_list(const char *path, uint64_t &n, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
                           std::string &err) {
    std::string p(path);
    const size_t dot = p.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : p.substr(dot);
    const bool mtx = ext == ".mtx" || ext == ".mmio", txt = ext == ".txt", tsv = ext == ".tsv";
    if (!mtx && !txt && !tsv) {
        err = "unsupported file type " + ext + ": only .tsv (GraphChallenge), .txt (SNAP), .mtx / .mmio (Matrix Market)";
        return false;
    }
    FILE *f = fopen(path, "r");
    if (!f) { err = "cannot open " + p; return false; }
    char line[4096];
    uint64_t maxid = 0, lineno = 0;
    bool size_line = mtx, any = false;
    while (fgets(line, sizeof(line), f)) {
        if (txt && lineno++ < 4) continue;  // SNAP header (skiprows=4)
        if (line[0] == '%' || line[0] == '#' || line[0] == '\n' || line[0] == '\r') continue;
        unsigned long long a, b;
        if (sscanf(line, "%llu %llu", &a, &b) != 2) continue;
        if (size_line) { size_line = false; continue; }  // rows cols nnz
        if (mtx) {
            if (a == 0 || b == 0) { err = "0 index in a 1-based Matrix Market file"; fclose(f); return false; }
            --a;
            --b;
        }
        if (a >= 0xFFFFFFFEull || b >= 0xFFFFFFFEull) { err = "vertex id too large"; fclose(f); return false; }
        src.push_back((uint32_t)a);
        dst.push_back((uint32_t)b);
        if (a > maxid) maxid = a;
        if (b > maxid) maxid = b;
        any = true;
    }
    fclose(f);
    n = any ? maxid + 1 : 0;
    return true;
}

}  // namespace lstc_text
