"""Write tests/golden/orient_undir2dir.json: orientation vectors produced by the
reference's own CPU tool preprocess_CPU/Undir2Dir.c, compiled from the
reference sources by ``make -C oracle ref`` into oracle/_ref/undir2dir.

The symmetric input CSRs are built here in plain Python (set semantics), so the
fixture depends on neither the oracle nor the product.  Undir2Dir leaves
row_ptr_Dir[i+1] unset for rows without neighbours (Undir2Dir.c:112-132), so
every input graph has no isolated vertices.

    make -C oracle ref && python tests/golden/make_orient_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import graphs as G  # noqa: E402

TOOL = os.path.join(ROOT, "oracle", "_ref", "undir2dir")


def sym_csr(n, src, dst):
    es = G.simple_edge_set(src, dst)
    verts = sorted({v for e in es for v in e})
    remap = {v: i for i, v in enumerate(verts)}  # order-preserving compaction
    adj = [[] for _ in verts]
    for a, b in es:
        adj[remap[a]].append(remap[b])
        adj[remap[b]].append(remap[a])
    off, col = [0], []
    for row in adj:
        col += sorted(row)
        off.append(len(col))
    return len(verts), off, col


def run_tool(n, off, col, workdir):
    deg = [off[i + 1] - off[i] for i in range(n)]
    inp, out = os.path.join(workdir, "in.txt"), os.path.join(workdir, "out.txt")
    with open(inp, "w") as f:
        f.write(f"{n}\n{' '.join(map(str, deg))}\n")
        f.write(f"{n} {n + 1} {len(col)} {n}\n")
        f.write(" ".join(str(i) for i in range(n)) + "\n")
        f.write(" ".join(map(str, off)) + "\n" + " ".join(map(str, col)) + "\n")
    subprocess.run([TOOL, inp, out, "1"], check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, cwd=workdir)
    tok = [int(x) for x in open(out).read().split()]
    t_ver, rps, pos = tok[0], tok[1], tok[2]
    assert t_ver == n and rps == n + 1
    return tok[3: 3 + rps], tok[3 + rps: 3 + rps + pos]


def main():
    cases = {
        "edge_filter_cu": G.toy_edge_filter_cu(),
        "K6": G.complete(6),
        "wheel7": G.wheel(7),
        "windmill4": G.windmill(4),
        "bipartite_3x4": G.complete_bipartite(3, 4),
        "random_n40_m120": G.random_multigraph(40, 120, seed=3),
        "random_n200_m900": G.random_multigraph(200, 900, seed=7),
    }
    out = {"_source": "outputs of the reference's preprocess_CPU/Undir2Dir.c built by oracle/Makefile "
                      "(target ref); inputs are symmetric CSRs without isolated vertices",
           "cases": {}}
    with tempfile.TemporaryDirectory() as wd:
        for name, (n, s, d) in cases.items():
            nn, off, col = sym_csr(n, s, d)
            doff, dcol = run_tool(nn, off, col, wd)
            out["cases"][name] = {"n": nn, "sym_off": off, "sym_col": col,
                                  "dir_off": doff, "dir_col": dcol,
                                  "triangles": G.brute_triangles(s, d)}
    with open(os.path.join(HERE, "orient_undir2dir.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
