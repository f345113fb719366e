"""Write rmat_counts.json: oracle triangle counts of the seeded RMAT workloads.

RMAT scale s, edge factor 16, a/b/c/d .57/.19/.19/.05, seed 1, cleaned and
degree-oriented by the oracle (oracle/lstc_oracle.c), counted by its merge
counter.  The GPU tests compare the device pipeline with these values at sizes
too large to recount on the CPU inside a test.

    python tests/golden/make_rmat_counts.py [scales...]     (default 16..23)
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    scales = [int(a) for a in sys.argv[1:]] or list(range(16, 24))
    path = os.path.join(HERE, "rmat_counts.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for s in scales:
        t0 = time.time()
        n, src, dst = O.gen_rmat(s, 16, seed=1)
        _, g = O.clean_orient(n, src, dst)
        del src, dst
        tri = O.count(g, threads=os.cpu_count() or 1)
        out[str(s)] = {"scale": s, "edge_factor": 16, "abcd": [0.57, 0.19, 0.19, 0.05], "seed": 1,
                       "n": int(g.n), "m": int(g.m), "triangles": int(tri)}
        print(f"s{s}: n={g.n} m={g.m} triangles={tri} ({time.time() - t0:.0f} s)", flush=True)
        json.dump(out, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
