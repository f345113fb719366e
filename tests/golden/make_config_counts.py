"""Write the oracle's golden counts of the full-size workloads that the GPU tests check.

    python tests/golden/make_config_counts.py halo      # -> halo_s23_k8.json
    python tests/golden/make_config_counts.py mixed     # -> mixed_counts.json
    python tests/golden/make_config_counts.py serving   # -> serving_counts.json

halo    RMAT scale 23, edge factor 16, a/b/c/d .57/.19/.19/.05, seed 1, cleaned and
        degree-oriented by the oracle; vertices assigned to 8 parts by the seeded
        random partition (seed 1); for every part the un-pruned halo subgraph
        (P u N(P), owned rows first) and its owned-row count, plus checksums of
        its CSR.  The parts sum to the unpartitioned count.
mixed   the 66 graphs of trianglecount_amd.workloads.mixed_specs(1): sizes and
        counts; the grid's count also equals twice its number of diagonals.
serving every request of trianglecount_amd.workloads.serving_stream(10000, 1000, 1):
        undirected edges after cleaning and the exact count.

All values come from oracle/ (the CPU restatement), run in this container.
"""
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from trianglecount_amd.workloads import mixed_specs, serving_stream  # noqa: E402

THREADS = os.cpu_count() or 1


def csr_checksums(off, col):
    """Size-independent fingerprints of a CSR: sum of offsets, sum of ids, and a
    position-weighted sum of ids (weight (i mod 1024) + 1), all exact in uint64."""
    off = np.asarray(off, np.uint64)
    col = np.asarray(col, np.uint64)
    w = (np.arange(len(col), dtype=np.uint64) & np.uint64(1023)) + np.uint64(1)
    return {"off_sum": int(off.sum(dtype=np.uint64)), "col_sum": int(col.sum(dtype=np.uint64)),
            "col_wsum": int((w * col).sum(dtype=np.uint64))}


def halo(scale=23, k=8, seed=1):
    t0 = time.time()
    n, src, dst = O.gen_rmat(scale, 16, seed=seed)
    sym, g = O.clean_orient(n, src, dst)
    del src, dst
    total = O.count(g, threads=THREADS)
    print(f"s{scale}: n={g.n} m={g.m} triangles={total} ({time.time() - t0:.0f} s)", flush=True)
    part = O.partition_random(g.n, k, seed=seed)

    def one(p):
        h = O.halo_extract(sym, part, p, mode=0)
        c = O.count(h.csr, n_owned=h.n_owned, threads=1)
        out = {"part": p, "n_local": h.csr.n, "n_owned": h.n_owned, "m_local": h.csr.m, "triangles": c}
        out.update(csr_checksums(h.csr.off, h.csr.col))
        print(f"  part {p}: {out} ({time.time() - t0:.0f} s)", flush=True)
        return out

    with cf.ThreadPoolExecutor(min(k, THREADS)) as ex:
        parts = list(ex.map(one, range(k)))
    assert sum(q["triangles"] for q in parts) == total, "parts do not sum to the whole"
    res = {"scale": scale, "edge_factor": 16, "abcd": [0.57, 0.19, 0.19, 0.05], "seed": seed, "k": k,
           "partition_seed": seed, "mode": 0, "n": g.n, "m": g.m, "triangles": total, "parts": parts}
    json.dump(res, open(os.path.join(HERE, f"halo_s{scale}_k{k}.json"), "w"), indent=1)


def _gen(spec):
    if spec.kind == "grid":
        return O.gen_grid(spec.a, spec.b, spec.p, spec.seed)
    if spec.kind == "chung_lu":
        return O.gen_chung_lu(spec.a, spec.gamma, spec.mean, spec.w_max, seed=spec.seed)
    if spec.kind == "er":
        return O.gen_er(spec.a, spec.p, seed=spec.seed)
    raise ValueError(spec.kind)


def mixed(seed=1):
    specs = mixed_specs(seed)

    def one(spec):
        n, s, d = _gen(spec)
        _, g = O.clean_orient(n, s, d)
        out = {"name": spec.name, "n": g.n, "m": g.m, "triangles": O.count(g, threads=1)}
        if spec.kind == "grid":
            out["diagonals"] = O.grid_num_diagonals(spec.a, spec.b, spec.p, spec.seed)
            assert out["triangles"] == 2 * out["diagonals"], "grid count is not twice its diagonals"
        return out

    t0 = time.time()
    with cf.ThreadPoolExecutor(THREADS) as ex:
        graphs = list(ex.map(one, specs))
    print(f"mixed: {len(graphs)} graphs, {sum(g['m'] for g in graphs)} edges, "
          f"{sum(g['triangles'] for g in graphs)} triangles ({time.time() - t0:.0f} s)", flush=True)
    json.dump({"seed": seed, "graphs": graphs}, open(os.path.join(HERE, "mixed_counts.json"), "w"), indent=1)


def serving(n_req=10000, rate=1000.0, seed=1):
    reqs = serving_stream(n_req, rate, seed)

    def one(r):
        if r.kind == "rmat":
            n, s, d = O.gen_rmat(r.a, 16, seed=r.seed, m_raw=r.m_raw)
        else:
            n, s, d = O.gen_er(r.a, 20.0, seed=r.seed, m_raw=r.m_raw)
        _, g = O.clean_orient(n, s, d)
        return g.m, O.count(g, threads=1)

    t0 = time.time()
    order = sorted(range(n_req), key=lambda i: -reqs[i].m_raw)  # largest first: the pool ends evenly
    res = [None] * n_req
    with cf.ThreadPoolExecutor(THREADS) as ex:
        for done, (i, v) in enumerate(zip(order, ex.map(lambda i: one(reqs[i]), order))):
            res[i] = v
            if done % 500 == 0:
                print(f"  {done}/{n_req} ({time.time() - t0:.0f} s)", flush=True)
    out = {"requests": n_req, "rate_per_s": rate, "seed": seed,
           "kind": "".join("R" if r.kind == "rmat" else "E" for r in reqs),
           "a": [r.a for r in reqs], "m_raw": [r.m_raw for r in reqs],
           "edges": [int(v[0]) for v in res], "triangles": [int(v[1]) for v in res]}
    print(f"serving: {n_req} requests, {sum(out['edges'])} edges ({time.time() - t0:.0f} s)", flush=True)
    json.dump(out, open(os.path.join(HERE, "serving_counts.json"), "w"), separators=(",", ":"))


if __name__ == "__main__":
    what = sys.argv[1:] or ["mixed", "halo", "serving"]
    for w in what:
        {"halo": halo, "mixed": mixed, "serving": serving}[w]()
